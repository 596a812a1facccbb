"""Long-range tail corrections of the pair potentials (host side).

Restates what ``PairPot.prepare_tailcorrections`` (yaff/pes/ext.pyx:665-687)
and ``pair_pot_tailcorr_cut`` / ``pair_pot_tailcorr_switch3``
(yaff/pes/pair_pot.c:178-254) compute, for LJ, MM3 and ExpRep:

    C0 = sum over ordered atom pairs (i, j), i == j included, not both below
         nlow and not both at or above nhigh, of the missing part of
         int r^2 V_ij(r) dr;
    C1 = -C0 - V_ij(rc) rc^3 / 3 summed alike (plain cutoff), or -C0 (Switch3);

ForcePartTailCorrection adds 2 pi C0 / V to the energy and 2 pi C1 / V to the
diagonal of the virial (yaff/pes/ff.py:1012-1057).

The missing part is int_rc^inf r^2 V dr (closed forms below) plus, with a
Switch3 of width w, int_{rc-w}^{rc} r^2 V(r) (1 - S(r)) dr with
S = 3x^2 - 2x^3, x = (rc - r)/w, evaluated by composite 40-point
Gauss-Legendre quadrature on subintervals that double in length from
rc - w (each subinterval lies at least its own length away from the pole of
the r^-n terms at r = 0, so every piece converges geometrically and the sum
is exact to rounding however wide the switch).  Pairs are grouped by their
per-atom parameters, so the cost is O(distinct parameter sets ^ 2), not
O(natom ^ 2).
"""
import numpy as np

__all__ = ['tail_integrals', 'prepare_tailcorrections']

_GL_X, _GL_W = np.polynomial.legendre.leggauss(40)
MM3_A = 1.84e5     # pair_pot.c (pair_fn_mm3): 1.84e5 exp(-12 r / sigma)
MM3_C6 = 2.25      # and 2.25 (sigma / r)^6


def _exp_tail(a, rc):
    """int_rc^inf r^2 exp(-a r) dr for a > 0."""
    return np.exp(-a*rc)*(rc*rc/a + 2.0*rc/(a*a) + 2.0/(a*a*a))


def _pot(kind, p, r):
    """V(r) of the pair parameters p (arrays broadcast against r)."""
    if kind == 'lj':
        x = (p['sigma']/r)**6
        return 4.0*p['epsilon']*x*(x - 1.0)
    if kind == 'mm3':
        return p['epsilon']*(MM3_A*np.exp(-12.0*r/p['sigma']) - p['r6scale']*MM3_C6*(p['sigma']/r)**6)
    if kind == 'exprep':
        return p['amp']*np.exp(-p['b']*r)
    raise ValueError(kind)


def _cut_tail(kind, p, rc):
    """int_rc^inf r^2 V(r) dr."""
    if kind == 'lj':
        s6 = p['sigma']**6
        return 4.0*p['epsilon']*(s6*s6/(9.0*rc**9) - s6/(3.0*rc**3))
    if kind == 'mm3':
        return p['epsilon']*(MM3_A*_exp_tail(12.0/p['sigma'], rc)
                             - p['r6scale']*MM3_C6*p['sigma']**6/(3.0*rc**3))
    if kind == 'exprep':
        b = p['b']
        ok = b != 0.0
        return np.where(ok, p['amp']*_exp_tail(np.where(ok, b, 1.0), rc), 0.0)
    raise ValueError(kind)


def tail_integrals(kind, p, rc, width=None):
    """(I0, I1) per pair parameter set: I0 the missing part of int r^2 V dr,
    I1 the virial integral (-I0 - V(rc) rc^3/3 with a plain cutoff, -I0
    with a Switch3 of the given width)."""
    i0 = _cut_tail(kind, p, rc)
    if width is None:
        return i0, -i0 - _pot(kind, p, rc)*rc**3/3.0
    r, wq = _switch_nodes(rc, width)
    x = (rc - r)/width
    keep = 1.0 - x*x*(3.0 - 2.0*x)
    pp = {k: np.asarray(v, float)[..., None] for k, v in p.items()}
    i0 = i0 + np.sum(wq*r*r*_pot(kind, pp, r)*keep, axis=-1)
    return i0, -i0


def _switch_nodes(rc, width):
    """Nodes and weights of the composite rule on [rc - width, rc]."""
    lo = rc - width
    if not lo > 0.0:
        raise ValueError('a Switch3 tail correction needs rcut > width')
    edges = [lo]
    while edges[-1] < rc and len(edges) < 64:
        edges.append(min(2.0*edges[-1], rc))
    edges[-1] = rc
    a, b = np.asarray(edges[:-1]), np.asarray(edges[1:])
    half = 0.5*(b - a)
    r = (0.5*(a + b))[:, None] + half[:, None]*_GL_X[None, :]
    return r.ravel(), (half[:, None]*_GL_W[None, :]).ravel()


def _groups(keys, natom, nlow, nhigh):
    """Distinct per-atom keys with their counts below nlow, in [nlow, nhigh)
    and at or above nhigh."""
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.ravel()
    idx = np.arange(natom)
    cnt = np.zeros((len(uniq), 3))
    for c, sel in enumerate((idx < nlow, (idx >= nlow) & (idx < nhigh), idx >= nhigh)):
        np.add.at(cnt[:, c], inv[sel], 1.0)
    return uniq, cnt


def prepare_tailcorrections(pair_pot, natom, nlow=0, nhigh=-1):
    """(C0, C1) of ext.pyx:665-687 for a PairPotLJ, PairPotMM3 or
    PairPotExpRep (plain cutoff or Switch3)."""
    from .pair_pot import Switch3
    if nhigh == -1:
        nhigh = natom
    if not (0 <= nlow <= nhigh <= natom):
        raise ValueError('need 0 <= nlow <= nhigh <= natom')
    tr = pair_pot.tr
    if tr is not None and not isinstance(tr, Switch3):
        raise NotImplementedError('Tail corrections not supported for %s' % tr)
    width = None if tr is None else tr.width
    kind = pair_pot.name
    if kind == 'lj':
        keys = np.stack([pair_pot.sigmas, pair_pot.epsilons], axis=1)
    elif kind == 'mm3':
        keys = np.stack([pair_pot.sigmas, pair_pot.epsilons, pair_pot.onlypaulis.astype(float)], axis=1)
    elif kind == 'exprep':
        keys = np.asarray(pair_pot.ffatype_ids, float).reshape(-1, 1)
    else:
        raise NotImplementedError('Tail corrections not supported for pair_%s' % kind)
    uniq, cnt = _groups(keys, natom, nlow, nhigh)
    a, b = np.meshgrid(np.arange(len(uniq)), np.arange(len(uniq)), indexing='ij')
    a, b = a.ravel(), b.ravel()
    tot = cnt.sum(axis=1)
    # ordered pairs of the two groups, less those both below nlow or both at/above nhigh
    w = tot[a]*tot[b] - cnt[a, 0]*cnt[b, 0] - cnt[a, 2]*cnt[b, 2]
    ua, ub = uniq[a], uniq[b]
    if kind == 'lj':  # pair_pot.c (pair_fn_lj): arithmetic sigma, geometric epsilon
        p = {'sigma': 0.5*(ua[:, 0] + ub[:, 0]), 'epsilon': np.sqrt(ua[:, 1]*ub[:, 1])}
    elif kind == 'mm3':  # pair_fn_mm3: sigma = sigma_i + sigma_j, onlypauli of either drops the r^-6 term
        p = {'sigma': ua[:, 0] + ub[:, 0], 'epsilon': np.sqrt(ua[:, 1]*ub[:, 1]),
             'r6scale': np.where(ua[:, 2] + ub[:, 2] == 0, 1.0, 0.0)}
    else:  # pair_fn_exprep: cross tables indexed by the two types
        ta, tb = ua[:, 0].astype(int), ub[:, 0].astype(int)
        p = {'amp': np.asarray(pair_pot.amp_cross)[ta, tb], 'b': np.asarray(pair_pot.b_cross)[ta, tb]}
    i0, i1 = tail_integrals(kind, p, pair_pot.rcut, width)
    return float(np.sum(w*i0)), float(np.sum(w*i1))
