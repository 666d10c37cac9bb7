"""Drop-in counterparts of piquasso's ``piquasso._math`` native modules."""
