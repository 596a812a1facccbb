"""Pair potentials and truncation schemes with Yaff's constructor signatures
(yaff/pes/ext.pyx:524-619 truncations, 627-2342 pair potentials).

The objects hold parameters; the arithmetic runs in libyaff_nb.so.  The
only host evaluation is ``Truncation.trunc_fn`` for a single distance, a
helper of the reference API (truncation.c:30-95).
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import ROW_DTYPE, SCALING_DTYPE

__all__ = ['Truncation', 'Switch3', 'Hammer', 'PairPot', 'PairPotLJ', 'PairPotMM3', 'PairPotExpRep', 'PairPotEI']


class Truncation(object):
    """Base class of the truncation schemes."""

    def trunc_fn(self, d, rcut):
        """(h, dh/dd) at distance d."""
        raise NotImplementedError


class Hammer(Truncation):
    """exp(tau/(d - rcut)) for d < rcut (truncation.c:30-41)."""

    def __init__(self, tau):
        self.tau = float(tau)

    def trunc_fn(self, d, rcut):
        if d < rcut:
            x = d - rcut
            h = math.exp(self.tau/x)
            return h, -h*self.tau/x/x
        return 0.0, 0.0

    def get_log(self):
        return 'hammer %s' % self.tau


class Switch3(Truncation):
    """3x^2 - 2x^3 with x = (rcut - d)/width in the last width before rcut
    (truncation.c:59-76)."""

    def __init__(self, width):
        self.width = float(width)

    def trunc_fn(self, d, rcut):
        if d < rcut:
            x = rcut - d
            if x > self.width:
                return 1.0, 0.0
            x /= self.width
            return (3 - 2*x)*x*x, -6*x*(1 - x)/self.width
        return 0.0, 0.0

    def get_log(self):
        return 'switch3 %s' % self.width


def _c_contig(a, dtype):
    a = np.asarray(a)
    if not a.flags['C_CONTIGUOUS']:
        raise AssertionError('array must be C contiguous')
    return a.astype(dtype, copy=False)


class PairPot(object):
    """Base class of the pair potentials (ext.pyx:627-744)."""
    name = None

    def __init__(self, rcut, tr):
        if tr is not None and not isinstance(tr, Truncation):
            raise TypeError('tr must be a Truncation instance')
        self._rcut = float(rcut)
        self.tr = tr

    @property
    def rcut(self):
        return self._rcut

    def get_truncation(self):
        return self.tr

    def prepare_tailcorrections(self, natom, nlow=0, nhigh=-1):
        """(C0, C1): the tail-correction sums of ext.pyx:665-687 (energy and
        virial integrals over the ordered atom pairs); LJ, MM3 and ExpRep
        with a plain cutoff or a Switch3."""
        from .tailcorr import prepare_tailcorrections
        return prepare_tailcorrections(self, natom, nlow, nhigh)

    def log(self):
        pass

    def _spec(self, stab):
        from .device import pot_spec
        return pot_spec(self, stab)

    def compute(self, neighs, stab, gpos, vtens, nneigh):
        """pair_pot_compute over explicit rows (ext.pyx:689-744, pair_pot.c:84-176),
        evaluated on the device.  gpos/vtens are added to; returns the energy.
        Each atom's gradient is summed over its rows in row order (the index
        is made here with a stable sort), so the result is deterministic."""
        from .engine import Engine
        if not torch.cuda.is_available():
            raise RuntimeError('PairPot.compute needs a HIP device')
        lib = _lib.load()
        neighs = np.ascontiguousarray(neighs[:nneigh]).astype(ROW_DTYPE, copy=False)
        stab = np.zeros(0, SCALING_DTYPE) if stab is None else np.ascontiguousarray(stab).astype(SCALING_DTYPE)
        if gpos is not None:
            assert gpos.flags['C_CONTIGUOUS'] and gpos.shape[1] == 3
        if vtens is not None:
            assert vtens.flags['C_CONTIGUOUS'] and vtens.shape == (3, 3)
        spec = self._spec(stab)
        natom = len(spec.p0)
        dev = torch.device('cuda')

        def up(a, dt=np.float64):
            if a is None:
                return None
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dt).copy()).to(dev)
        rows_d = torch.from_numpy(np.frombuffer(neighs.tobytes() or b'\0'*72, np.uint8).copy()).to(dev)
        stab_d = torch.from_numpy(np.frombuffer(stab.tobytes() or b'\0'*32, np.uint8).copy()).to(dev)
        p0, p1, p2 = up(spec.p0), up(spec.p1), up(spec.p2)
        cross = None
        pot = _lib.YnbPot(spec.kind, spec.trunc, spec.rcut, spec.trunc_par, spec.alpha, spec.dielectric, 0, 0, 0, 0)
        if spec.amp_cross is not None:
            pot.nffatype = spec.amp_cross.shape[0]
            cross = up(np.concatenate([np.ravel(spec.amp_cross), np.ravel(spec.b_cross)]))
        derivs = gpos is not None or vtens is not None
        g_d = ref0_d = ref_d = rowh = None
        if gpos is not None:
            g_d = torch.zeros((natom, 3), dtype=torch.float64, device=dev)
            ref0, ref = _lib.row_refs(neighs, natom)
            ref0_d, ref_d = up(ref0, np.int32), up(ref if len(ref) else np.zeros(1, np.int32), np.int32)
            rowh = torch.empty(3*max(len(neighs), 1), dtype=torch.float64, device=dev)
        nblk = max((len(neighs) + 255)//256, 1)
        partial = torch.empty(8*nblk, dtype=torch.float64, device=dev)
        out = torch.zeros(10, dtype=torch.float64, device=dev)
        ptr = lambda t: None if t is None else t.data_ptr()
        _lib.check(lib.ynb_pair_rows(rows_d.data_ptr(), len(neighs), stab_d.data_ptr(), len(stab), pot, ptr(p0),
                                     ptr(p1), ptr(p2), ptr(cross), 1 if derivs else 0, natom, ptr(ref0_d),
                                     ptr(ref_d), ptr(g_d), ptr(rowh), partial.data_ptr(), out.data_ptr(),
                                     Engine.stream()))
        o = out.cpu().numpy()
        if gpos is not None:
            gpos += g_d.cpu().numpy()
        if vtens is not None:
            vtens += o[1:].reshape(3, 3)
        return float(o[0])


class PairPotLJ(PairPot):
    """4 eps ((sigma/d)^12 - (sigma/d)^6), sigma = (s_a + s_b)/2,
    eps = sqrt(e_a e_b) (pair_pot.c:277-301, ext.pyx:747-819)."""
    name = 'lj'

    def __init__(self, sigmas, epsilons, rcut, tr=None):
        PairPot.__init__(self, rcut, tr)
        self.sigmas = _c_contig(sigmas, np.float64)
        self.epsilons = _c_contig(epsilons, np.float64)
        assert self.sigmas.shape == self.epsilons.shape


class PairPotMM3(PairPot):
    """eps (1.84e5 exp(-12 d/sigma) - 2.25 (sigma/d)^6), sigma = s_a + s_b
    (pair_pot.c:321-352, ext.pyx:822-908)."""
    name = 'mm3'

    def __init__(self, sigmas, epsilons, onlypaulis, rcut, tr=None):
        PairPot.__init__(self, rcut, tr)
        self.sigmas = _c_contig(sigmas, np.float64)
        self.epsilons = _c_contig(epsilons, np.float64)
        self.onlypaulis = _c_contig(onlypaulis, np.int32)
        assert self.sigmas.shape == self.epsilons.shape == self.onlypaulis.shape


class PairPotExpRep(PairPot):
    """A_ij exp(-B_ij d) with optional mixing rules (pair_pot.c:473-489,
    ext.pyx:1038-1195)."""
    name = 'exprep'

    def __init__(self, ffatype_ids, amp_cross, b_cross, rcut, tr=None, amps=None, amp_mix=-1, amp_mix_coeff=-1,
                 bs=None, b_mix=-1, b_mix_coeff=-1):
        PairPot.__init__(self, rcut, tr)
        ffatype_ids = _c_contig(ffatype_ids, np.int64)
        amp_cross = np.array(_c_contig(amp_cross, np.float64))
        b_cross = np.array(_c_contig(b_cross, np.float64))
        nff = amp_cross.shape[0]
        assert amp_cross.shape == (nff, nff) and b_cross.shape == (nff, nff)
        assert ffatype_ids.min() >= 0 and ffatype_ids.max() < nff
        if amps is not None:
            assert len(amps) == nff and amp_mix in (0, 1) and 0 <= amp_mix_coeff <= 1
            self._init_amp_cross(nff, amp_cross, amps, amp_mix, amp_mix_coeff)
        assert (amp_cross == amp_cross.T).all()
        if bs is not None:
            assert len(bs) == nff and b_mix in (0, 1) and 0 <= b_mix_coeff <= 1
            self._init_b_cross(nff, b_cross, bs, b_mix, b_mix_coeff, amps)
        assert (b_cross == b_cross.T).all()
        self.ffatype_ids = ffatype_ids
        self.amp_cross = amp_cross
        self.b_cross = b_cross

    @staticmethod
    def _init_amp_cross(nff, amp_cross, amps, amp_mix, coeff):
        """ext.pyx:1146-1158"""
        for i0 in range(nff):
            for i1 in range(i0 + 1):
                if amp_cross[i0, i1] == 0.0:
                    if amp_mix == 0:
                        amp_cross[i0, i1] = np.sqrt(amps[i0]*amps[i1])
                    elif amps[i0] == 0.0 or amps[i1] == 0.0:
                        pass
                    else:
                        amp = (np.log(amps[i0]) + np.log(amps[i1]))/2
                        amp *= 1 - coeff*abs(np.log(amps[i0]/amps[i1]))
                        amp_cross[i0, i1] = np.exp(amp)
                    amp_cross[i1, i0] = amp_cross[i0, i1]

    @staticmethod
    def _init_b_cross(nff, b_cross, bs, b_mix, coeff, amps):
        """ext.pyx:1160-1172"""
        for i0 in range(nff):
            for i1 in range(i0 + 1):
                if b_cross[i0, i1] == 0.0:
                    if b_mix == 0:
                        b_cross[i0, i1] = (bs[i0] + bs[i1])/2
                    elif amps[i0] == 0.0 or amps[i1] == 0.0:
                        pass
                    else:
                        b = (bs[i0] + bs[i1])/2
                        b *= 1 - coeff*abs(np.log(amps[i0]/amps[i1]))
                        b_cross[i0, i1] = b
                    b_cross[i1, i0] = b_cross[i0, i1]


class PairPotEI(PairPot):
    """Real-space Ewald (or bare, alpha = 0) Coulomb between point or Gaussian
    charges (pair_pot.c:711-751, ext.pyx:1770-1857)."""
    name = 'ei'

    def __init__(self, charges, alpha, rcut, tr=None, dielectric=1.0, radii=None):
        PairPot.__init__(self, rcut, tr)
        self.charges = _c_contig(charges, np.float64)
        self.alpha = float(alpha)
        self.dielectric = float(dielectric)
        self.radii = np.zeros(self.charges.shape) if radii is None else _c_contig(radii, np.float64)
