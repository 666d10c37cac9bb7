#!/bin/sh
# Compile one kernel object: hipcc "$@" with kernel-resource-usage remarks.
# The remarks (registers, scratch, occupancy per kernel) go to <out>.res;
# other diagnostics are shown, minus hipcc's host-pass note about the
# packed-FP32 feature flag.  A kernel that spills to scratch is reported.
# usage: hipcc_kernel.sh <hipcc> <out.o> <args...>
HIPCC="$1"; OUT="$2"; shift 2
LOG="$OUT.log"
"$HIPCC" -Rpass-analysis=kernel-resource-usage "$@" -o "$OUT.tmp" 2> "$LOG"
st=$?
grep -E "Function Name|VGPRs:|AGPRs:|ScratchSize|Occupancy" "$LOG" \
  | sed -e 's/^.*remark: *//' -e 's/ \[-Rpass.*//' > "$OUT.res"
grep -v -e "packed-fp32-ops' is not a recognized feature" -e "remark:" "$LOG" \
  | grep -v -E "^ *[0-9]* *[|]|^ *[|]|^ *\^|^[0-9]+ (warning|remark)s? generated" >&2
rm -f "$LOG"
if grep -q "ScratchSize \[bytes/lane\]: [1-9]" "$OUT.res"; then
  echo "  WARNING: $OUT spills to scratch (see $OUT.res)" >&2
fi
[ $st -eq 0 ] && mv "$OUT.tmp" "$OUT" && echo "  HIPCC $OUT"
exit $st
