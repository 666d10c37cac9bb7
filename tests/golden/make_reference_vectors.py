"""Write tests/golden/reference_vectors.json: the known-answer vectors that the
reference's own test-suite holds for the permanent path, transcribed as data
(inputs and expected outputs, each with the reference file:line it comes from).

Run:  python tests/golden/make_reference_vectors.py
"""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

U3 = [
    [0.62270314 + 0.55117657j, -0.0258677 - 0.07171713j, 0.09597446 - 0.54168404j],
    [0.34756795 - 0.29444499j, -0.43514701 + 0.18975153j, 0.71929752 + 0.22304973j],
    [-0.24500645 - 0.20227626j, -0.45222962 - 0.75121057j, 0.06995606 - 0.3540245j],
]
U6A = [
    [0.62113733 - 0.01959968j, -0.15627468 - 0.0772489j, -0.2705819 - 0.18997122j,
     0.26504798 - 0.30838768j, 0.03372169 - 0.11154586j, 0.15278476 + 0.52137824j],
    [-0.1776024 - 0.21000195j, 0.18950753 + 0.20741494j, -0.15537846 + 0.19161071j,
     0.07400899 - 0.37578572j, -0.44458249 - 0.0047501j, -0.62212719 + 0.23055313j],
    [-0.05572001 - 0.20287464j, 0.22359337 + 0.30693557j, -0.13719319 + 0.23245719j,
     0.1102451 + 0.02659467j, 0.81942653 + 0.04327346j, -0.17215559 + 0.15114287j],
    [-0.24319645 - 0.44143551j, -0.50022937 - 0.08513718j, 0.07671116 - 0.05858231j,
     0.0679656 + 0.52109972j, -0.0482276 - 0.12736588j, -0.11768435 + 0.41307881j],
    [-0.29469977 - 0.20027018j, 0.22135149 - 0.02983563j, -0.18587346 - 0.83950064j,
     -0.21606625 - 0.14975436j, 0.11702974 - 0.02297493j, -0.01552763 + 0.01646485j],
    [-0.29741767 + 0.15644426j, -0.61959257 - 0.23497653j, 0.07397837 + 0.05367843j,
     -0.05838964 - 0.57132173j, 0.28736069 - 0.00798998j, -0.13763068 - 0.09058005j],
]
U4 = [
    [0.50142122 - 0.15131566j, 0.0265964 - 0.67076793j, -0.48706586 - 0.15221542j,
     0.1148288 - 0.0381445j],
    [-0.30504953 + 0.02783877j, -0.33442264 + 0.27130083j, -0.52279136 - 0.6644152j,
     -0.04684954 + 0.06143215j],
    [0.20279518 + 0.57890235j, -0.47584393 - 0.16777288j, 0.08168919 + 0.10673884j,
     -0.34434342 + 0.48221603j],
    [0.31022502 - 0.39919511j, 0.21927998 + 0.24751057j, -0.06266598 - 0.05334444j,
     -0.78447544 + 0.11350836j],
]
U6B = [
    [0.00456448 - 0.37857979j, 0.0795763 + 0.10465352j, 0.03365652 + 0.41152925j,
     0.05555206 + 0.18738511j, -0.39720048 + 0.1311295j, 0.40848797 + 0.537455j],
    [0.00901246 + 0.19215191j, 0.4609619 + 0.13507775j, -0.47801241 + 0.19203936j,
     0.23140442 - 0.62025569j, 0.00333535 + 0.14816728j, 0.0783928 + 0.02267907j],
    [0.43945864 - 0.04054749j, 0.00377335 - 0.30622153j, 0.22692424 - 0.43506116j,
     0.08710784 - 0.32486556j, -0.58818325 + 0.08888961j, 0.04507539 - 0.0406845j],
    [-0.29415733 + 0.17675044j, 0.33184779 + 0.39750559j, 0.34950358 - 0.26546825j,
     -0.45096811 - 0.10690984j, -0.12192328 + 0.20376462j, -0.237482 + 0.30640823j],
    [-0.02410776 + 0.30238262j, -0.46322751 + 0.20107186j, -0.09817621 - 0.15398772j,
     -0.29490926 - 0.28320581j, 0.02993443 - 0.3052611j, 0.5788743 + 0.15254606j],
    [0.53706284 + 0.35589712j, -0.26260555 + 0.25708476j, 0.23762169 + 0.20229218j,
     0.16489218 + 0.03524027j, 0.24701479 + 0.48757485j, -0.06188047 + 0.14647274j],
]
INT10 = [
    [0.0, 0.0, -1.0, -1.0, -1.0, 0.0, -1.0, 0.0, 0.0, -1.0],
    [-1.0, -1.0, -1.0, -1.0, 0.0, 0.0, -1.0, -1.0, -1.0, -1.0],
    [-1.0, 0.0, -1.0, -1.0, -1.0, 0.0, 0.0, -1.0, 0.0, -1.0],
    [-1.0, 0.0, -1.0, -1.0, -1.0, -1.0, 0.0, -1.0, -1.0, -1.0],
    [-1.0, 0.0, 0.0, -1.0, -1.0, -1.0, -1.0, -1.0, 0.0, -1.0],
    [0.0, -1.0, 0.0, 0.0, -1.0, -1.0, -1.0, -1.0, 0.0, -1.0],
    [-1.0, 0.0, 0.0, 0.0, -1.0, -1.0, -1.0, -1.0, 0.0, -1.0],
    [0.0, 0.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 0.0, 0.0],
    [0.0, 0.0, 0.0, -1.0, -1.0, -1.0, 0.0, 0.0, -1.0, -1.0],
    [0.0, 0.0, 0.0, 0.0, -1.0, 0.0, -1.0, -1.0, -1.0, 0.0],
]
ASYM23 = [
    [0.909394 + 0.264435j, 0.00450261 + 0.0188079j, 0.316704 + 0.0490014j],
    [-0.109435 - 0.117915j, 0.35198 + 0.812257j, 0.312276 + 0.304881j],
]

GRAD6A = [
    [-0.05371236 + 2.66353718e-02j, -0.04582236 - 8.18820862e-02j, 0.03960814 + 5.80689287e-02j,
     -0.02045293 + 1.38801402e-01j, -0.0149418 - 4.19752512e-02j, -0.05229726 - 5.69852986e-02j],
    [-0.11103513 + 3.88283318e-02j, -0.14391923 + 1.05461540e-02j, 0.00725286 - 8.25634953e-03j,
     0.05221953 + 2.01817507e-02j, 0.01545186 - 4.29582156e-03j, -0.00398563 - 7.07909136e-02j],
    [0.04208331 + 4.16242008e-02j, -0.00862918 - 5.05756669e-03j, 0.02234116 + 1.42226190e-02j,
     -0.02339448 - 3.14244737e-03j, 0.05196793 + 5.23028802e-02j, 0.00677792 + 9.11863479e-02j],
    [-0.06956478 + 2.20011711e-02j, 0.09569006 + 4.02943688e-02j, -0.05576987 + 4.91285771e-02j,
     -0.00468075 - 6.13113076e-02j, 0.02943685 + 4.46223167e-02j, 0.05577522 - 2.92991471e-02j],
    [0.011846 + 2.46933770e-03j, 0.0153745 + 4.67444609e-02j, -0.04146899 + 2.13750603e-02j,
     0.04391381 + 4.23307063e-02j, -0.0189607 - 2.91139206e-02j, -0.05233688 + 1.27779869e-04j],
    [-0.03119874 + 7.55431598e-02j, -0.15592711 - 2.44391509e-02j, 0.01571884 - 4.49747533e-02j,
     -0.01909996 - 9.40045056e-02j, -0.02374948 - 7.11344685e-02j, 0.04924467 - 3.22028064e-02j],
]
GRAD6B = [
    [0.00613756 - 0.05596387j, 0.05478057 + 0.04177092j, 0.01846764 + 0.11305476j,
     0.02605859 + 0.01649895j, -0.01002297 - 0.03639016j, 0.02551702 - 0.04844018j],
    [0.00808861 - 0.01126793j, 0.00079495 + 0.05701643j, -0.02790181 - 0.01987287j,
     0.00725814 - 0.0567719j, 0.0763283 + 0.01962546j, 0.02445184 + 0.05709078j],
    [-0.06986495 - 0.00897261j, -0.05773915 - 0.08317558j, -0.02051542 + 0.01104616j,
     -0.02175885 - 0.04162514j, -0.08400035 - 0.01669359j, 0.00111028 - 0.05966386j],
    [0.05612247 + 0.00807638j, -0.02331525 + 0.00396841j, -0.00055769 - 0.02280161j,
     -0.04036997 - 0.04793161j, -0.01883466 + 0.01268112j, 0.04200545 - 0.02340915j],
    [0.02595464 - 0.0743598j, -0.04747572 + 0.01271623j, -0.0620618 + 0.00135199j,
     -0.05118836 - 0.00165811j, 0.01362084 - 0.06675524j, -0.09843736 + 0.05753861j],
    [0.04273109 + 0.00110463j, 0.00367212 - 0.00144348j, 0.01679765 - 0.00517546j,
     -0.00579187 - 0.00950378j, -0.02463916 + 0.11107093j, -0.03805067 - 0.05460198j],
]
GRAD4 = [
    [0.0, 0.0, -3.18426548 - 1.54032741j, -0.09908472 - 0.37279947j],
    [0.0, 0.0, 0.0, 0.0],
    [0.0, 0.0, 0.0, 0.0],
    [0.0, 0.0, 0.54425976 + 0.15950196j, -0.49015741 - 0.62882923j],
]

P3 = [
    [1, 0, 0],
    [0, -0.54687158 + 0.07993182j, 0.32028583 - 0.76938896j],
    [0, 0.78696803 + 0.27426941j, 0.42419041 - 0.35428818j],
]
P3C = [
    [-0.25022099 + 0.32110177j, -0.69426529 - 0.49960543j, -0.28233272 + 0.15153042j],
    [-0.69028768 + 0.23351228j, 0.14839865 + 0.49185272j, -0.43153658 - 0.13714903j],
    [0.41073351 + 0.36681879j, -0.06655274 + 0.00442722j, -0.22146243 - 0.80202711j],
]


def cm(x):
    a = np.asarray(x, dtype=np.complex128)
    return {"re": a.real.tolist(), "im": a.imag.tolist()}


def main():
    ones = lambda k: [1] * k  # noqa: E731
    perm_cases = [
        # src: tests/_math/test_permanent.py
        dict(id="trivial_real", src="tests/_math/test_permanent.py:24-29", matrix=cm([[4.2]]),
             real_input=True, rows=[1], cols=[1], expected=cm(4.2), rtol=1e-5),
        dict(id="zero_input", src="tests/_math/test_permanent.py:32-55", matrix=cm(U3),
             rows=[0, 0, 0], cols=[0, 0, 0], expected=cm(1.0), rtol=1e-5),
        dict(id="no_repetition_6x6", src="tests/_math/test_permanent.py:58-117", matrix=cm(U6A),
             rows=ones(6), cols=ones(6),
             expected=cm(0.022227325527358795 + 0.01807052573717885j), rtol=1e-5),
        dict(id="asymmetric_2x2", src="tests/_math/test_permanent.py:120-126",
             matrix=cm(np.array([[1, 1j], [1, -1j]]) / np.sqrt(2)), rows=[2, 0], cols=[0, 2],
             expected=cm(-1.0), rtol=1e-5),
        dict(id="mult_4x4", src="tests/_math/test_permanent.py:129-165", matrix=cm(U4),
             rows=[3, 0, 0, 1], cols=[0, 0, 3, 1],
             expected=cm(0.4302957973670928 + 0.3986355418194044j), rtol=1e-5),
        dict(id="mult_6x6", src="tests/_math/test_permanent.py:168-228", matrix=cm(U6B),
             rows=[1, 1, 0, 3, 2, 2], cols=[2, 1, 3, 0, 1, 2],
             expected=cm(0.13160241373727416 + 0.36535625772184577j), rtol=1e-5),
        dict(id="integer_10x10_zero", src="tests/_math/test_permanent.py:252-282", matrix=cm(INT10),
             rows=[0, 0, 0, 0, 0, 3, 4, 0, 3, 0], cols=[1, 0, 0, 0, 0, 0, 1, 7, 0, 1],
             expected=cm(0.0), rtol=0.0, exact=True),
        dict(id="asymmetric_2x3", src="tests/_math/test_permanent.py:285-301", matrix=cm(ASYM23),
             rows=[2, 1], cols=[1, 1, 1], expected=None, check="equals_expanded"),
        # src: tests/jax_extensions/test_permanent_unit.py
        dict(id="empty_0x0", src="tests/jax_extensions/test_permanent_unit.py:58-61",
             matrix={"re": [], "im": []}, shape=[0, 0], rows=[], cols=[], expected=cm(1.0),
             rtol=1e-5),
        dict(id="single_zero", src="tests/jax_extensions/test_permanent_unit.py:64-67",
             matrix=cm([[0.0]]), rows=[1], cols=[1], expected=cm(0.0), rtol=0.0, exact=True),
        dict(id="negative_entries", src="tests/jax_extensions/test_permanent_unit.py:78-82",
             matrix=cm([[1, -2], [-3, 4]]), rows=[1, 1], cols=[1, 1], expected=cm(10.0),
             rtol=0.0, exact=True),
        dict(id="large_values", src="tests/jax_extensions/test_permanent_unit.py:85-91",
             matrix=cm(np.full((3, 3), 1e10)), rows=[1, 1, 1], cols=[1, 1, 1],
             expected=cm(6e30), rtol=1e-10),
        dict(id="high_repetition", src="tests/jax_extensions/test_permanent_unit.py:94-102",
             matrix=cm([[2, 3], [4, 5]]), rows=[3, 0], cols=[2, 1], expected=cm(72.0),
             rtol=0.0, exact=True),
        dict(id="identity_8", src="tests/jax_extensions/test_permanent_unit.py:111-115",
             matrix=cm(np.eye(8)), rows=ones(8), cols=ones(8), expected=cm(1.0), rtol=0.0,
             exact=True),
        dict(id="conjugate_symmetry", src="tests/jax_extensions/test_permanent_unit.py:125-130",
             matrix=cm([[1 + 1j, 2 - 1j], [3 + 0j, 4 + 2j]]), rows=[1, 1], cols=[1, 1],
             expected=None, check="conjugate_symmetry"),
    ]
    grad_cases = [
        # src: tests/jax_extensions/test_permanent_grad.py (holomorphic d perm / d A)
        dict(id="grad_3x3_exact", src="tests/jax_extensions/test_permanent_grad.py:43-56",
             matrix=cm([[1, 2, 3], [4, 5, 6], [7, 8, 9]]), rows=ones(3), cols=ones(3),
             expected=cm([[93, 78, 67], [42, 30, 22], [27, 18, 13]]), rtol=0.0, exact=True),
        dict(id="grad_identity", src="tests/jax_extensions/test_permanent_grad.py:70-74",
             matrix=cm(np.eye(3)), rows=ones(3), cols=ones(3), expected=cm(np.eye(3)), rtol=1e-5),
        dict(id="grad_single", src="tests/jax_extensions/test_permanent_grad.py:88-92",
             matrix=cm([[2.0]]), rows=[1], cols=[1], expected=cm([[1.0]]), rtol=1e-5),
        dict(id="grad_zero", src="tests/jax_extensions/test_permanent_grad.py:106-110",
             matrix=cm(np.zeros((2, 2))), rows=[1, 1], cols=[1, 1], expected=cm(np.zeros((2, 2))),
             rtol=1e-5),
        dict(id="grad_ones", src="tests/jax_extensions/test_permanent_grad.py:124-128",
             matrix=cm(np.ones((2, 2))), rows=[1, 1], cols=[1, 1], expected=cm(np.ones((2, 2))),
             rtol=1e-5),
        dict(id="grad_zero_input", src="tests/jax_extensions/test_permanent_grad.py:149-172",
             matrix=cm(U3), rows=[0, 0, 0], cols=[0, 0, 0], expected=cm(np.zeros((3, 3))),
             rtol=1e-5),
        dict(id="grad_6x6_no_repetition", src="tests/jax_extensions/test_permanent_grad.py:175-285",
             matrix=cm(U6A), rows=ones(6), cols=ones(6), expected=cm(GRAD6A), rtol=1e-5),
        dict(id="grad_4x4_mult", src="tests/jax_extensions/test_permanent_grad.py:301-354",
             matrix=cm(U4), rows=[3, 0, 0, 1], cols=[0, 0, 3, 1], expected=cm(GRAD4), rtol=1e-5),
        dict(id="grad_6x6", src="tests/jax_extensions/test_permanent_grad.py:368-479",
             matrix=cm(U6B), rows=ones(6), cols=ones(6), expected=cm(GRAD6B), rtol=1e-5),
        dict(id="grad_2x2_rows20", src="tests/jax_extensions/test_permanent_grad.py:495-514",
             matrix=cm([[0.5, -0.8660254], [0.8660254, 0.5]]), rows=[2, 0], cols=[2, 0],
             expected=cm([[2.0, 0.0], [0.0, 0.0]]), rtol=1e-5),
    ]
    prob_cases = [
        # src: tests/_simulators/passive/*: p = |perm(U; out, in)|^2 / (prod out! prod in!)
        dict(id="detection_probability", src="tests/_simulators/passive/test_preparations.py:231-249",
             matrix=cm(P3), input=[1, 1, 0], output=[1, 1, 0], expected=0.30545762086020883),
        dict(id="detection_probability_complex",
             src="tests/_simulators/passive/test_preparations.py:252-282",
             matrix=cm(P3C), input=[0, 1, 2], output=[2, 1, 0], expected=0.038483056956364094),
        dict(id="identity_probability", src="tests/_simulators/passive/test_state.py:71-92",
             matrix=cm(np.eye(2)), input=[1, 1], output=[1, 1], expected=1.0),
        dict(id="beamsplitter_probability", src="tests/_simulators/passive/test_state.py:94-117",
             matrix=cm([[np.cos(np.pi / 5), -np.sin(np.pi / 5)], [np.sin(np.pi / 5), np.cos(np.pi / 5)]]),
             input=[1, 1], output=[2, 0], expected=float(np.sin(2 * np.pi / 5) ** 2 / 2)),
    ]
    out = {"permanent": perm_cases, "gradient": grad_cases, "probability": prob_cases}
    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
