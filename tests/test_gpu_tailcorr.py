"""Tail corrections with the device pair parts, as the reference checks them
(yaff/pes/test/test_tailcorr.py:40-99, 173-196): for a 64-atom LJ fluid the
energy and pressure with tail corrections stay at their rcut -> infinity
limits (extrapolated from E(rc) = E_inf + c/rc^3 of the uncorrected runs) for
every cutoff from 12 to 32 A, plain and with a Switch3, over all atoms and
with a frame excluded (nlow)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

ANG = 1.889726133921252
KCALMOL = 0.0015936015865590541
MPA = 1e6/2.9421015697e13  # Hartree/bohr^3 per MPa


def _fluid(natom=64, l0=3.0*ANG, seed=4):
    import yaff_amd as ya
    n = int(np.ceil(natom**(1.0/3.0)))
    g = np.arange(n)*l0
    pos = np.stack(np.meshgrid(g, g, g, indexing='ij'), axis=-1).reshape(-1, 3)[:natom]
    pos = pos + np.random.default_rng(seed).normal(0.0, 0.01*l0, pos.shape)
    return ya.System(np.ones(natom, int), pos, ffatypes=['X'], ffatype_ids=np.zeros(natom, int), bonds=[],
                     rvecs=np.eye(3)*l0*n)


@pytest.mark.parametrize('nlow', [0, 59])
@pytest.mark.parametrize('width', [None, 3.0*ANG])
def test_lj_fluid_converges(nlow, width):
    import yaff_amd as ya
    ref = _fluid()
    natom = ref.natom
    rng = np.random.default_rng(7)
    sigmas = 2.1*ANG + rng.normal(0.0, 0.01*ANG, natom)
    epsilons = 0.15*KCALMOL + rng.normal(0.0, 0.001*KCALMOL, natom)
    rcuts = np.linspace(12.0, 32.0, 11)*ANG
    data = np.zeros((len(rcuts), 4))
    for i, rc in enumerate(rcuts):
        system = ya.System(ref.numbers, ref.pos, ffatype_ids=ref.ffatype_ids, bonds=[], rvecs=ref.cell.rvecs)
        nlist = ya.NeighborList(system, nlow=nlow)
        tr = None if width is None else ya.Switch3(width)
        part = ya.ForcePartPair(system, nlist, ya.Scalings(system), ya.PairPotLJ(sigmas, epsilons, rc, tr))
        tail = ya.ForcePartTailCorrection(system, part, nlow=nlow)
        ff0 = ya.ForceField(system, [part], nlist=nlist)
        ff1 = ya.ForceField(system, [part, tail], nlist=nlist)
        vol = system.cell.volume
        v0, v1 = np.zeros((3, 3)), np.zeros((3, 3))
        e0 = ff0.compute(vtens=v0)
        e1 = ff1.compute(vtens=v1)
        data[i] = [e0, e1, np.trace(v0)/3.0/vol, np.trace(v1)/3.0/vol]
    eref = np.polyfit(1.0/rcuts**3, data[:, 0], 1)[1]
    pref = np.polyfit(1.0/rcuts**3, data[:, 2], 1)[1]
    assert np.abs(data[:, 1] - eref).max() < 0.05*KCALMOL
    assert np.abs(data[:, 3] - pref).max() < 0.1*MPA
    # and the corrections matter: without them the energy at 12 A is off by more
    assert abs(data[0, 0] - eref) > 2*abs(data[0, 1] - eref)
