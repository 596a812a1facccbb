"""Force-field parts with Yaff's protocol (yaff/pes/ff.py:65-669).

``ForcePart.compute(gpos, vtens)`` has the reference semantics: each part
computes into a private zeroed buffer, the energy and buffers are checked for
NaN, and the results are *added* to the caller's arrays
(ff.py:117-168).  Behind it, all ForcePartPair objects of a system are
evaluated by one fused device launch per position state, and the Ewald parts
by the device kernels of libyaff_nb.so.
"""
import numpy as np

from .device import SystemDevice

__all__ = ['ForcePart', 'ForceField', 'ForcePartPair', 'ForcePartEwaldReciprocal', 'ForcePartEwaldCorrection',
           'ForcePartEwaldNeutralizing', 'ForcePartTailCorrection']


def check_nlow_nhigh(system, nlow, nhigh):
    """ff.py check of the nlow/nhigh pair (yaff/pes/ff.py:1265-1277)."""
    if nlow < 0:
        raise ValueError('nlow must be a positive number, received %d.' % nlow)
    if nhigh == -1:
        nhigh = system.natom
    if nhigh < nlow:
        raise ValueError('nhigh must not be smaller than nlow, received %d.' % nhigh)
    return nlow, nhigh


class ForcePart(object):
    """ff.py:65-172"""

    def __init__(self, name, system):
        self.name = name
        self.energy = 0.0
        self.gpos = np.zeros((system.natom, 3), float)
        self.vtens = np.zeros((3, 3), float)
        self.clear()

    def clear(self):
        self.energy = np.nan
        self.gpos[:] = np.nan
        self.vtens[:] = np.nan

    def update_rvecs(self, rvecs):
        self.clear()

    def update_pos(self, pos):
        self.clear()

    def compute(self, gpos=None, vtens=None):
        if gpos is None:
            my_gpos = None
        else:
            my_gpos = self.gpos
            my_gpos[:] = 0.0
        if vtens is None:
            my_vtens = None
        else:
            my_vtens = self.vtens
            my_vtens[:] = 0.0
        self.energy = self._internal_compute(my_gpos, my_vtens)
        if np.isnan(self.energy):
            raise ValueError('The energy is not-a-number (nan).')
        if gpos is not None:
            if np.isnan(my_gpos).any():
                raise ValueError('Some gpos element(s) is/are not-a-number (nan).')
            gpos += my_gpos
        if vtens is not None:
            if np.isnan(my_vtens).any():
                raise ValueError('Some vtens element(s) is/are not-a-number (nan).')
            vtens += my_vtens
        return self.energy

    def _internal_compute(self, gpos, vtens):
        raise NotImplementedError


class ForceField(ForcePart):
    """ff.py:175-273"""

    def __init__(self, system, parts, nlist=None):
        ForcePart.__init__(self, 'all', system)
        self.system = system
        self.parts = []
        self.nlist = nlist
        self.needs_nlist_update = nlist is not None
        for part in parts:
            self.add_part(part)

    @classmethod
    def generate(cls, system, parameters, skip_other_sections=False, **kwargs):
        """ff.py:217-251: a force field from a Yaff parameter file (a
        filename, a list of them or a Parameters instance) with the
        nonbonded generators LJ, MM3, EXPREP and FIXQ; keywords as
        generator.FFArgs (rcut, tr, alpha_scale, gcut_scale, skin, smooth_ei,
        reci_ei, nlow, nhigh, tailcorrections).  Sections of Yaff's other
        generators raise NotImplementedError unless skip_other_sections."""
        from .generator import FFArgs, apply_generators
        from .parameters import Parameters
        if system.ffatype_ids is None:
            raise ValueError('The generators needs ffatype_ids in the system object.')
        if not isinstance(parameters, Parameters):
            parameters = Parameters.from_file(parameters)
        ff_args = FFArgs(**kwargs)
        apply_generators(system, parameters, ff_args, skip_other_sections)
        return cls(system, ff_args.parts, ff_args.nlist)

    def add_part(self, part):
        self.parts.append(part)
        name = 'part_%s' % part.name
        if name in self.__dict__:
            raise ValueError('The part %s occurs twice in the force field.' % name)
        self.__dict__[name] = part

    def update_rvecs(self, rvecs):
        ForcePart.update_rvecs(self, rvecs)
        self.system.cell.update_rvecs(rvecs)
        SystemDevice.of(self.system).invalidate()
        if self.nlist is not None:
            self.nlist.update_rmax()
            self.needs_nlist_update = True

    def update_pos(self, pos):
        ForcePart.update_pos(self, pos)
        self.system.pos[:] = pos
        if self.nlist is not None:
            self.needs_nlist_update = True

    def _internal_compute(self, gpos, vtens):
        if self.needs_nlist_update:
            self.nlist.update()
            self.needs_nlist_update = False
        else:
            SystemDevice.of(self.system).upload_positions()
        return sum([part.compute(gpos, vtens) for part in self.parts])


class ForcePartPair(ForcePart):
    """ff.py:276-328.  All pair parts of a system share one fused launch."""

    def __init__(self, system, nlist, scalings, pair_pot):
        ForcePart.__init__(self, 'pair_%s' % pair_pot.name, system)
        self.system = system
        self.nlist = nlist
        self.scalings = scalings
        self.pair_pot = pair_pot
        self.nlist.request_rcut(pair_pot.rcut)
        self.slot = SystemDevice.of(system).add_pair(nlist, pair_pot, scalings.stab)

    def _internal_compute(self, gpos, vtens):
        e, g, v = SystemDevice.of(self.system).pair_results()[self.slot]
        if gpos is not None:
            gpos[:] = g
        if vtens is not None:
            vtens[:] = v
        return e


class ForcePartEwaldReciprocal(ForcePart):
    """ff.py:331-401"""

    def __init__(self, system, alpha, gcut=0.35, dielectric=1.0, nlow=0, nhigh=-1):
        ForcePart.__init__(self, 'ewald_reci', system)
        if not system.cell.nvec == 3:
            raise TypeError('The system must have a 3D periodic cell.')
        if system.charges is None:
            raise ValueError('The system does not have charges.')
        self.system = system
        self.alpha = alpha
        self.gcut = gcut
        self.dielectric = dielectric
        self.update_gmax()
        self.nlow, self.nhigh = check_nlow_nhigh(system, nlow, nhigh)
        SystemDevice.of(system).set_ewald(self)

    def update_gmax(self):
        """ff.py:384-389"""
        self.gmax = np.ceil(self.gcut/self.system.cell.gspacings - 0.5).astype(int)

    def update_rvecs(self, rvecs):
        ForcePart.update_rvecs(self, rvecs)
        self.update_gmax()
        SystemDevice.of(self.system).invalidate()

    def _internal_compute(self, gpos, vtens):
        dev = SystemDevice.of(self.system)
        if dev.nlist is None or dev.list_token < 0:
            dev.upload_positions()
        e, g, v = dev.ewald_results()
        if gpos is not None:
            gpos[:] = g
        if vtens is not None:
            vtens[:] = v
        return e


class ForcePartEwaldCorrection(ForcePart):
    """ff.py:472-532"""

    def __init__(self, system, alpha, scalings, dielectric=1.0, nlow=0, nhigh=-1):
        ForcePart.__init__(self, 'ewald_cor', system)
        if not system.cell.nvec == 3:
            raise TypeError('The system must have a 3D periodic cell')
        if system.charges is None:
            raise ValueError('The system does not have charges.')
        self.system = system
        self.alpha = alpha
        self.dielectric = dielectric
        self.nlow, self.nhigh = check_nlow_nhigh(system, nlow, nhigh)
        self.scalings = scalings

    def _internal_compute(self, gpos, vtens):
        dev = SystemDevice.of(self.system)
        if dev.nlist is None or dev.list_token < 0:
            dev.upload_positions()
        e, g, v = dev.ewald_corr(self.alpha, self.scalings.stab, self.dielectric, self.nlow, self.nhigh)
        if gpos is not None:
            gpos[:] = g
        if vtens is not None:
            vtens[:] = v
        return e


class ForcePartEwaldNeutralizing(ForcePart):
    """ff.py:590-669: O(1) background term of a charged cell (host)."""

    def __init__(self, system, alpha, dielectric=1.0, nlow=0, nhigh=-1, fluctuating_charges=False):
        ForcePart.__init__(self, 'ewald_neut', system)
        if not system.cell.nvec == 3:
            raise TypeError('The system must have a 3D periodic cell')
        if system.charges is None:
            raise ValueError('The system does not have charges.')
        self.system = system
        self.alpha = alpha
        self.dielectric = dielectric
        self.nlow, self.nhigh = check_nlow_nhigh(system, nlow, nhigh)
        self.fluctuating_charges = fluctuating_charges
        if not fluctuating_charges:
            self.prefactor = self._fac()*np.pi/(2.0*self.dielectric)

    def _fac(self):
        q = self.system.charges
        fac = q[:].sum()**2/self.alpha**2
        fac -= q[:self.nlow].sum()**2/self.alpha**2
        fac -= q[self.nhigh:].sum()**2/self.alpha**2
        r = self.system.radii
        if r is not None:
            fac -= q.sum()*np.sum(q*r**2)
            fac += q[:self.nlow].sum()*np.sum(q[:self.nlow]*r[:self.nlow]**2)
            fac += q[self.nhigh:].sum()*np.sum(q[self.nhigh:]*r[self.nhigh:]**2)
        return fac

    def _internal_compute(self, gpos, vtens):
        if not self.fluctuating_charges:
            fac = self.prefactor/self.system.cell.volume
        else:
            fac = self._fac()*np.pi/(2.0*self.system.cell.volume*self.dielectric)
        if vtens is not None:
            vtens.ravel()[::4] -= fac
        return fac


class ForcePartTailCorrection(ForcePart):
    """ff.py:1012-1057: long-range corrections of a pair part's energy and
    virial for a homogeneous system (host, O(1) per evaluation after set-up;
    the integrals are made once by PairPot.prepare_tailcorrections)."""

    def __init__(self, system, part_pair, nlow=0, nhigh=-1):
        if system.cell.nvec != 3:
            raise ValueError('Tail corrections can only be applied to 3D periodic systems')
        if part_pair.name in ['pair_ei', 'pair_eidip']:
            raise ValueError('Tail corrections are divergent for %s' % part_pair.name)
        ForcePart.__init__(self, 'tailcorr_%s' % part_pair.name, system)
        self.nlow, self.nhigh = check_nlow_nhigh(system, nlow, nhigh)
        self.ecorr, self.wcorr = part_pair.pair_pot.prepare_tailcorrections(system.natom, self.nlow, self.nhigh)
        self.system = system

    def _internal_compute(self, gpos, vtens):
        if vtens is not None:
            w = 2.0*np.pi*self.wcorr/self.system.cell.volume
            vtens[0, 0] += w
            vtens[1, 1] += w
            vtens[2, 2] += w
        return 2.0*np.pi*self.ecorr/self.system.cell.volume
