"""SCF-side terms: non-XC energy, exact exchange (filled in with the J/K kernels)."""


def non_xc_energy(molecule, rdm1):
    raise NotImplementedError("non-XC energy needs the packed-ERI J kernel")


def exact_exchange_energy(molecule, rdm1, a0):
    raise NotImplementedError("exact exchange needs the packed-ERI K kernel")
