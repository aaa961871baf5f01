"""MolecularSurfaceCalculator on MI355X: the reference's calculator API
(cms-cuda/_core.py:762-2527, _gpu.py:310-516) over the HIP library.

The host side keeps the atoms added with AddMolecule / add_binder_and_target
as NumPy arrays; the first Calc* call moves them to the GPU (torch tensors
as device memory, the current torch stream as the HIP stream), generates
the surface once in libcms_hip.so and caches it, exactly like
``_ensure_surfaces_generated``.  Results come back as Python floats and
float64 NumPy arrays.  ``calc.run`` mirrors the reference's run state
(results counters, dots, probes, atoms) and is copied from the device only
when read.
"""

import ctypes
import weakref

import numpy as np

from . import _lib, tuning

ATTEN_BLOCKER = 1
ATTEN_2 = 2
ATTEN_BURIED_FLAGGED = 5
ATTEN_6 = 6
MAX_SUBDIV = 100


def _torch():
    import torch
    return torch


def _device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("cms_cuda needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Settings:
    """Calculator settings with the reference defaults (_core.py:766-773)."""

    def __init__(self):
        self.rp = 1.7
        self.density = 15.0
        self.band = 1.5
        self.sep = 8.0
        self.weight = 0.5
        self.binwidth_dist = 0.02
        self.binwidth_norm = 0.02

    def to_c(self):
        return _lib.cms_settings(self.rp, self.density, self.band, self.sep, self.weight,
                                 self.binwidth_dist, self.binwidth_norm)


# ------------------------------------------------------------ run-state mirror

class ResultsSurface:
    def __init__(self):
        self.d_mean = 0.0
        self.d_median = 0.0
        self.s_mean = 0.0
        self.s_median = 0.0
        self.nAtoms = 0
        self.nBuriedAtoms = 0
        self.nBlockedAtoms = 0
        self.nAllDots = 0
        self.nTrimmedDots = 0
        self.nBuriedDots = 0
        self.nAccessibleDots = 0
        self.trimmedArea = 0.0


class ResultsDots:
    def __init__(self):
        self.convex = 0
        self.concave = 0
        self.toroidal = 0


class RESULTS:
    def __init__(self):
        self.sc = 0.0
        self.area = 0.0
        self.distance = 0.0
        self.perimeter = 0.0
        self.nAtoms = 0
        self.surface = [ResultsSurface(), ResultsSurface(), ResultsSurface()]
        self.dots = ResultsDots()
        self.valid = 0


class DotArray:
    """Host copy of one molecule's dots (fields as in _core.DotArray)."""

    def __init__(self, coor_xyz, outnml_xyz, area, buried, type_, atom_idx):
        self.coor_xyz = coor_xyz
        self.outnml_xyz = outnml_xyz
        self.area = area
        self.buried = buried
        self.type_ = type_
        self.atom_idx = atom_idx

    @classmethod
    def empty(cls):
        return cls(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.int8),
                   np.zeros(0, np.int8), np.zeros(0, np.int32))

    def __len__(self):
        return len(self.area)

    def __bool__(self):
        return len(self.area) > 0


class ProbeArray:
    def __init__(self, atoms, height, point, alt):
        self.atom_idx_0 = atoms[:, 0].copy()
        self.atom_idx_1 = atoms[:, 1].copy()
        self.atom_idx_2 = atoms[:, 2].copy()
        self.height = height
        self.point_xyz = point
        self.alt_xyz = alt

    def __len__(self):
        return len(self.height)

    def __bool__(self):
        return len(self.height) > 0


class NeighborArray:
    """Host copy of neighbour or buried rows (the fields of the reference's
    SimpleNeighborArray, _core.py:596-611): natom (N, K) int32 padded with
    -1, radius (N, K) and xyz (N, K, 3) padded with NaN, nneighbors (N);
    K = the longest row (at least 1)."""

    def __init__(self, nneighbors, natom, xyz, radius):
        n, k = natom.shape
        self.nneighbors = nneighbors.astype(np.int32)
        self.natom = natom.astype(np.int32)
        self.xyz = np.full((n, k, 3), np.nan)
        self.radius = np.full((n, k), np.nan)
        m = self.natom >= 0
        self.xyz[m] = xyz[self.natom[m]]
        self.radius[m] = radius[self.natom[m]]

    def __len__(self):
        return len(self.nneighbors)

    def __bool__(self):
        return len(self.nneighbors) > 0


class AtomArray:
    def __init__(self, xyz, radius, density, molecule):
        n = len(radius)
        self.xyz = xyz
        self.radius = radius
        self.density = density
        self.molecule = molecule
        self.natom = np.arange(n, dtype=np.int32)
        self.nresidue = np.zeros(n, dtype=np.int32)
        self.atten = np.zeros(n, dtype=np.int8)
        self.access = np.zeros(n, dtype=np.int8)
        self.atom_name = [""] * n
        self.residue_name = [""] * n

    def __len__(self):
        return len(self.radius)

    def __bool__(self):
        return len(self.radius) > 0


class RunState:
    """Lazy mirror of the reference's ``calc.run``."""

    def __init__(self, calc):
        self._calc = weakref.proxy(calc)
        self.results = RESULTS()
        self.surfaces_generated = False
        self.surfaces_generated_all_atoms = None
        self._dots = None
        self._probes = None
        self._atoms = None
        self._trimmed = None
        self._rows = None
        self._toroid = None

    @property
    def radmax(self):
        r = self._calc._radii_all()
        return float(r.max()) if len(r) else 0.0

    @property
    def dots(self):
        if self._dots is None:
            self._dots = self._calc._copy_dots()
        return self._dots

    @property
    def probes(self):
        if self._probes is None:
            self._probes = self._calc._copy_probes()
        return self._probes

    @property
    def atoms(self):
        if self._atoms is None:
            self._atoms = self._calc._copy_atoms()
        return self._atoms

    @property
    def trimmed_dots(self):
        """Buried dots kept by the SC trim, per molecule (empty until
        CalcLoadedSC; _core.py:828-838)."""
        if self._trimmed is None:
            self._trimmed = self._calc._copy_trimmed()
        return self._trimmed

    @property
    def neighbor_array(self):
        if self._rows is None:
            self._rows = self._calc._copy_rows()
        return self._rows[0]

    @property
    def buried_array(self):
        if self._rows is None:
            self._rows = self._calc._copy_rows()
        return self._rows[1]

    @property
    def toroid_queue(self):
        """List of (atom 1, atom 2, uij, tij, rij, between) in the reference's
        order (_core.py:1419-1431)."""
        if self._toroid is None:
            self._toroid = self._calc._copy_toroid_queue()
        return self._toroid


# ------------------------------------------------------------ calculator

class MolecularSurfaceCalculator:
    """Drop-in for cms_cuda.MolecularSurfaceCalculator / ...GPU, running on
    the HIP library."""

    def __init__(self):
        self.settings = Settings()
        self._ctx = None
        self._fin = None
        self._full = None
        self.reset()

    # ---------------------------------------------------------- state
    def reset(self):
        self._release()
        self._xyz = [np.zeros((0, 3)), np.zeros((0, 3))]
        self._rad = [np.zeros(0), np.zeros(0)]
        self._dens = [np.zeros(0), np.zeros(0)]
        self._mode = None
        self._cms = None
        self._sc = None
        self._maxcms = None
        self._dev = None
        self.run = RunState(self)

    def _acquire(self, device, stream):
        # the finalizer holds the context, so garbage collection of the
        # calculator returns it to the pool instead of destroying it
        self._ctx = _lib.acquire_context(device, stream)
        self._fin = weakref.finalize(self, _lib.release_context, self._ctx)
        # the calculator has no backward pass: no dot descriptors
        self._ctx("cms_set_keep_descriptors", 0)

    def _release(self):
        fin = getattr(self, "_fin", None)
        if fin is not None:
            fin()
        self._fin = None
        self._ctx = None
        full = getattr(self, "_full", None)
        if full is not None:
            full.close()
        self._full = None

    def _radii_all(self):
        return np.concatenate(self._rad)

    def add_binder_and_target(self, binder_xyz, binder_radii, target_xyz, target_radii):
        """_core.py:918-921."""
        self.reset()
        self.AddMolecule(0, binder_xyz, binder_radii)
        self.AddMolecule(1, target_xyz, target_radii)

    def AddMolecule(self, molecule, xyz, radii):
        """Append atoms to molecule 0 or 1 (_core.py:923-943).  Atoms with
        radius <= 0 are kept so per-atom outputs stay aligned."""
        m = 1 if molecule == 1 else 0
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        radii = np.asarray(radii, dtype=np.float64).reshape(-1)
        if len(xyz) != len(radii):
            raise ValueError("xyz and radii lengths differ")
        self._xyz[m] = np.concatenate([self._xyz[m], xyz])
        self._rad[m] = np.concatenate([self._rad[m], radii])
        self._dens[m] = np.concatenate([self._dens[m], np.full(len(radii), float(self.settings.density))])
        self.run.results.surface[m].nAtoms += len(radii)
        self.run.results.nAtoms += len(radii)

    # ---------------------------------------------------------- surface
    def _ensure_surfaces_generated(self, all_atoms=False):
        """Generate the surface once; a later call in the other all_atoms mode
        raises (_core.py:865-909)."""
        if self._mode is not None:
            if self._mode != all_atoms:
                raise RuntimeError(
                    "This MolecularSurfaceCalculator instance already generated its "
                    f"molecular surface with all_atoms={self._mode}, "
                    f"but was asked to reuse it with all_atoms={all_atoms}. These use "
                    "incompatible attention-number assignments (CalcLoadedMaxPossibleCMS's "
                    "all_atoms=True vs. CalcLoaded()/CalcLoadedSC()'s all_atoms=False), so "
                    "the cached surface can't be safely reused for both. Call self.reset() "
                    "and re-add your atoms, or use a fresh MolecularSurfaceCalculator().")
            return
        n0, n1 = len(self._rad[0]), len(self._rad[1])
        assert n0 + n1 > 0
        torch = _torch()
        dev = _device()
        stream = torch.cuda.current_stream(dev).cuda_stream
        if self._ctx is None:
            self._acquire(dev.index, stream)
        xyz = torch.from_numpy(np.ascontiguousarray(np.concatenate(self._xyz))).to(dev)
        rad = torch.from_numpy(np.ascontiguousarray(self._radii_all())).to(dev)
        dens = torch.from_numpy(np.ascontiguousarray(np.concatenate(self._dens))).to(dev)
        self._dev = dev
        nb = (ctypes.c_int32 * 1)(n0)
        nt = (ctypes.c_int32 * 1)(n1)
        s = self.settings.to_c()
        tuning.apply(self._ctx, tuning.choice(n0 + n1))
        self._ctx("cms_load", 1, nb, nt, _ptr(xyz), _ptr(rad), _ptr(dens), ctypes.byref(s))
        try:
            self._ctx("cms_generate", 1 if all_atoms else 0, 0)
        except _lib.CmsError as e:
            if e.code == _lib.CMS_ERR_COINCIDENT:
                i = ctypes.c_int32()
                j = ctypes.c_int32()
                self._ctx("cms_coincident", 0, ctypes.byref(i), ctypes.byref(j))
                raise RuntimeError("Coincident atoms: %d:: == %d::" % (i.value, j.value)) from None
            raise
        self._mode = all_atoms
        self.run.surfaces_generated = True
        self.run.surfaces_generated_all_atoms = all_atoms
        self._fill_counts(all_atoms)

    def _fill_counts(self, all_atoms):
        out = np.zeros(_lib.NCOUNTS, dtype=np.int64)
        self._ctx("cms_counts", ctypes.c_void_p(out.ctypes.data))
        res = self.run.results
        res.dots.convex = int(out[4])
        res.dots.toroidal = int(out[5])
        res.dots.concave = int(out[6])
        n = [len(self._rad[0]), len(self._rad[1])]
        blocked = [int(out[10]), int(out[11])]
        for m in (0, 1):
            surf = res.surface[m]
            if all_atoms:
                surf.nBuriedAtoms += n[m]
            else:
                surf.nBlockedAtoms += blocked[m]
                surf.nBuriedAtoms += n[m] - blocked[m]
        self._counts = out

    def generate_molecular_surfaces(self, all_atoms=False):
        self._ensure_surfaces_generated(all_atoms)

    # ---------------------------------------------------------- readouts
    def CalcLoaded(self):
        """Target-side CMS: (cms, per-target-atom CMS).  _core.py:791-797."""
        self.run.results.valid = 0
        self._ensure_surfaces_generated(all_atoms=False)
        return self.calc_contact_molecular_surface(target_side=True)

    def CalcLoadedMaxPossibleCMS(self):
        """Surface area of molecule 0 (_core.py:799-805)."""
        self.run.results.valid = 0
        self._ensure_surfaces_generated(all_atoms=True)
        return self.calc_max_possible_contact_molecular_surface(target_side=True)

    def _compute_cms(self, precision=_lib.PREC_FP64):
        torch = _torch()
        n = len(self._rad[0]) + len(self._rad[1])
        per_atom = torch.empty(max(n, 1), dtype=torch.float64, device=self._dev)
        totals = torch.empty(2, dtype=torch.float64, device=self._dev)
        self._ctx("cms_contact_ms", int(precision), _ptr(per_atom), _ptr(totals))
        return per_atom[:n].cpu().numpy(), totals.cpu().numpy()

    def calc_contact_molecular_surface(self, target_side=True):
        """(total CMS, per-atom CMS of the query molecule): query = target
        (molecule 1) when target_side, else the binder.  _core.py:946-1004."""
        n0, n1 = len(self._rad[0]), len(self._rad[1])
        nq = n1 if target_side else n0
        # like the reference, read whatever surface exists (none: zeros)
        if self._mode is None:
            return 0.0, np.zeros(nq, dtype=np.float64)
        if self._cms is None:
            self._cms = self._compute_cms()
        per_atom, totals = self._cms
        if target_side:
            return float(totals[1]), per_atom[n0:n0 + n1].copy()
        return float(totals[0]), per_atom[:n0].copy()

    def calc_max_possible_contact_molecular_surface(self, target_side=True):
        """(total area, per-atom area) of molecule 0.  _core.py:1007-1034."""
        n0 = len(self._rad[0])
        if self._mode is None:
            return 0.0, np.zeros(n0, dtype=np.float64)
        if self._maxcms is None:
            torch = _torch()
            n = n0 + len(self._rad[1])
            per_atom = torch.empty(max(n, 1), dtype=torch.float64, device=self._dev)
            totals = torch.empty(1, dtype=torch.float64, device=self._dev)
            self._ctx("cms_max_contact_ms", _ptr(per_atom), _ptr(totals))
            self._maxcms = (per_atom[:n].cpu().numpy(), totals.cpu().numpy())
        per_atom, totals = self._maxcms
        return float(totals[0]), per_atom[:n0].copy()

    def CalcLoadedSC(self):
        """Lawrence & Colman SC: (sc, sc_int_area, median_dist).
        _core.py:807-863."""
        self.run.results.valid = 0
        self._ensure_surfaces_generated(all_atoms=False)
        if self._sc is None:
            torch = _torch()
            res = torch.empty(_lib.SC_NRES, dtype=torch.float64, device=self._dev)
            self._ctx("cms_shape_complementarity", _lib.PREC_FP64, _ptr(res))
            self._sc = res.cpu().numpy()
        r = self._sc
        out = self.run.results
        for m in (0, 1):
            side = r[8 + 8 * m: 16 + 8 * m]
            surf = out.surface[m]
            surf.d_mean, surf.d_median, surf.s_mean, surf.s_median = map(float, side[:4])
            surf.trimmedArea = float(side[4])
            surf.nTrimmedDots = int(side[5])
            surf.nAllDots = int(side[6])
        s0, s1, s2 = out.surface
        s2.d_mean = (s0.d_mean + s1.d_mean) / 2
        s2.d_median = (s0.d_median + s1.d_median) / 2
        s2.s_mean = (s0.s_mean + s1.s_mean) / 2
        s2.s_median = (s0.s_median + s1.s_median) / 2
        s2.nAllDots = s0.nAllDots + s1.nAllDots
        s2.nTrimmedDots = s0.nTrimmedDots + s1.nTrimmedDots
        s2.trimmedArea = s0.trimmedArea + s1.trimmedArea
        out.sc = s2.s_median
        out.distance = s2.d_median
        out.area = s2.trimmedArea
        out.valid = 1
        return out.sc, out.area, out.distance

    # ---------------------------------------------------------- introspection
    def _copy_dots(self):
        if self._mode is None:
            return [DotArray.empty(), DotArray.empty()]
        torch = _torch()
        out = []
        for m in (0, 1):
            n = ctypes.c_int64()
            self._ctx("cms_dot_count", 0, m, ctypes.byref(n))
            n = n.value
            if n == 0:
                out.append(DotArray.empty())
                continue
            dev = self._dev
            coor = torch.empty((n, 3), dtype=torch.float64, device=dev)
            nml = torch.empty((n, 3), dtype=torch.float64, device=dev)
            area = torch.empty(n, dtype=torch.float64, device=dev)
            bur = torch.empty(n, dtype=torch.int8, device=dev)
            typ = torch.empty(n, dtype=torch.int8, device=dev)
            atom = torch.empty(n, dtype=torch.int32, device=dev)
            self._ctx("cms_copy_dots", 0, m, _ptr(coor), _ptr(nml), _ptr(area), _ptr(bur), _ptr(typ), _ptr(atom))
            out.append(DotArray(coor.cpu().numpy(), nml.cpu().numpy(), area.cpu().numpy(),
                                bur.cpu().numpy(), typ.cpu().numpy(), atom.cpu().numpy()))
        return out

    def _copy_probes(self):
        if self._mode is None:
            z = np.zeros((0, 3))
            return ProbeArray(np.zeros((0, 3), np.int32), np.zeros(0), z, z)
        torch = _torch()
        n = ctypes.c_int64()
        self._ctx("cms_probe_count", 0, ctypes.byref(n))
        n = n.value
        dev = self._dev
        atoms = torch.empty((max(n, 1), 3), dtype=torch.int32, device=dev)
        h = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        p = torch.empty((max(n, 1), 3), dtype=torch.float64, device=dev)
        a = torch.empty((max(n, 1), 3), dtype=torch.float64, device=dev)
        if n:
            self._ctx("cms_copy_probes", 0, _ptr(atoms), _ptr(h), _ptr(p), _ptr(a))
        return ProbeArray(atoms[:n].cpu().numpy(), h[:n].cpu().numpy(), p[:n].cpu().numpy(),
                          a[:n].cpu().numpy())

    def _full_state(self):
        """A context holding a full-state generation of the same inputs and
        mode (rows for every atom, identical dots): the reference's per-atom
        state and neighbour tables cover every atom, the cached surface only
        the interface region."""
        if self._full is None:
            torch = _torch()
            xyz = np.concatenate(self._xyz)
            full = _lib.Context(self._dev.index, torch.cuda.current_stream(self._dev).cuda_stream)
            try:
                d_xyz = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.float64)).to(self._dev)
                d_rad = torch.from_numpy(np.ascontiguousarray(self._radii_all(), dtype=np.float64)).to(self._dev)
                d_den = torch.from_numpy(np.ascontiguousarray(np.concatenate(self._dens), dtype=np.float64)).to(self._dev)
                nb = (ctypes.c_int32 * 1)(len(self._rad[0]))
                nt = (ctypes.c_int32 * 1)(len(self._rad[1]))
                st = self.settings.to_c()
                full("cms_load", 1, nb, nt, _ptr(d_xyz), _ptr(d_rad), _ptr(d_den), ctypes.byref(st))
                full("cms_generate", 1 if self._mode else 0, 1)
                torch.cuda.current_stream(self._dev).synchronize()
            except Exception:
                full.close()
                raise
            self._full = full
        return self._full

    def _copy_atoms(self):
        xyz = np.concatenate(self._xyz)
        rad = self._radii_all()
        mol = np.concatenate([np.zeros(len(self._rad[0]), np.int8), np.ones(len(self._rad[1]), np.int8)])
        arr = AtomArray(xyz, rad, np.concatenate(self._dens), mol)
        if self._mode is not None and len(rad):
            torch = _torch()
            att = torch.empty(len(rad), dtype=torch.int8, device=self._dev)
            acc = torch.empty(len(rad), dtype=torch.int8, device=self._dev)
            self._full_state()("cms_copy_atom_state", _ptr(att), _ptr(acc))
            arr.atten = att.cpu().numpy()
            arr.access = acc.cpu().numpy()
        return arr

    def _copy_rows(self):
        """(neighbor_array, buried_array) of every atom, or (None, None)
        before a surface exists."""
        if self._mode is None:
            return None, None
        torch = _torch()
        n = len(self._radii_all())
        xyz = np.concatenate(self._xyz)
        rad = self._radii_all()
        out = []
        full = self._full_state()
        for which in (0, 1):
            cnt = torch.zeros(max(n, 1), dtype=torch.int32, device=self._dev)
            full("cms_copy_rows", 0, which, _ptr(cnt), None, 0)
            counts = cnt[:n].cpu().numpy()
            k = max(int(counts.max()) if n else 0, 1)
            nat = torch.empty((max(n, 1), k), dtype=torch.int32, device=self._dev)
            full("cms_copy_rows", 0, which, None, _ptr(nat), k)
            out.append(NeighborArray(counts, nat[:n].cpu().numpy(), xyz, rad))
        return tuple(out)

    def _copy_trimmed(self):
        if self._mode is None or self._sc is None:
            return [DotArray.empty(), DotArray.empty()]
        torch = _torch()
        out = []
        for m in (0, 1):
            n = ctypes.c_int64()
            self._ctx("cms_trimmed_count", 0, m, ctypes.byref(n))
            n = n.value
            if n == 0:
                out.append(DotArray.empty())
                continue
            dev = self._dev
            coor = torch.empty((n, 3), dtype=torch.float64, device=dev)
            nml = torch.empty((n, 3), dtype=torch.float64, device=dev)
            area = torch.empty(n, dtype=torch.float64, device=dev)
            bur = torch.empty(n, dtype=torch.int8, device=dev)
            typ = torch.empty(n, dtype=torch.int8, device=dev)
            atom = torch.empty(n, dtype=torch.int32, device=dev)
            self._ctx("cms_copy_trimmed_dots", 0, m, _ptr(coor), _ptr(nml), _ptr(area), _ptr(bur), _ptr(typ),
                      _ptr(atom))
            out.append(DotArray(coor.cpu().numpy(), nml.cpu().numpy(), area.cpu().numpy(),
                                bur.cpu().numpy(), typ.cpu().numpy(), atom.cpu().numpy()))
        return out

    def _copy_toroid_queue(self):
        if self._mode is None:
            return []
        torch = _torch()
        n = ctypes.c_int64()
        self._ctx("cms_toroid_count", 0, ctypes.byref(n))
        n = n.value
        if n == 0:
            return []
        dev = self._dev
        at = torch.empty((n, 2), dtype=torch.int32, device=dev)
        u = torch.empty((n, 3), dtype=torch.float64, device=dev)
        t = torch.empty((n, 3), dtype=torch.float64, device=dev)
        r = torch.empty(n, dtype=torch.float64, device=dev)
        b = torch.empty(n, dtype=torch.int8, device=dev)
        self._ctx("cms_copy_toroid_queue", 0, _ptr(at), _ptr(u), _ptr(t), _ptr(r), _ptr(b))
        at, u, t, r, b = (x.cpu().numpy() for x in (at, u, t, r, b))
        return [(int(at[i, 0]), int(at[i, 1]), u[i], t[i], float(r[i]), bool(b[i])) for i in range(n)]
