"""Python mirror of Alamo's elastic operator and solver API over the C-ABI of
libalamo_elastic.so (include/alamo_elastic.h).

  ElasticOperator  ~ Operator::Elastic<SYM>       src/Operator/Elastic.H:21-190
  Linear           ~ Solver::Nonlocal::Linear     src/Solver/Nonlocal/Linear.H:18-389
  Newton           ~ Solver::Nonlocal::Newton<T>  src/Solver/Nonlocal/Newton.H:14-654

Fields are torch float64 CUDA tensors shaped (comp, nz, ny, nx) with node
index i fastest.  Every operation runs in the HIP library on the current
torch stream; there is no CPU fallback: without the built library or a GPU
the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import bc as _bc
from . import model as _model

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libalamo_elastic.so")

ISOTROPIC, MAJORMINOR, DIAGONAL = _model.ISOTROPIC, _model.MAJORMINOR, _model.DIAGONAL
SOA, AOS = 0, 1
PREC_FP64, PREC_FP32_IR = 0, 1
CYCLE_F, CYCLE_V = 0, 1
OUTER_MLMG, OUTER_FGMRES = 0, 1
NOT_CONVERGED = 1


class AlamoError(RuntimeError):
    """A negative status from the library (Util::Abort in the reference)."""


class SolverError(AlamoError):
    """MLMG did not converge (amrex::Abort / exception in the reference)."""


class Opts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("cycle", ctypes.c_int), ("bottom_max_iter", ctypes.c_int), ("final_smooth", ctypes.c_int),
                ("omega", ctypes.c_double), ("bottom_tol_rel", ctypes.c_double), ("precision", ctypes.c_int),
                ("outer", ctypes.c_int), ("restart", ctypes.c_int), ("verbose", ctypes.c_int),
                ("bottom_solver", ctypes.c_int), ("learned", ctypes.c_int), ("smoother", ctypes.c_int),
                ("max_fmg_iter", ctypes.c_int), ("fixed_iter", ctypes.c_int)]


class Info(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int), ("res0", ctypes.c_double),
                ("rhsnorm", ctypes.c_double), ("res", ctypes.c_double), ("seconds", ctypes.c_double),
                ("learned_steps", ctypes.c_int)]


EXPORTS = [
    "alamo_el_last_error", "alamo_el_default_opts", "alamo_el_create", "alamo_el_destroy",
    "alamo_el_num_levels", "alamo_el_level_nodes", "alamo_el_set_bc", "alamo_el_set_periodic", "alamo_el_set_uniform",
    "alamo_el_set_omega", "alamo_el_set_model", "alamo_el_prepare", "alamo_el_prepare_fp32", "alamo_el_fapply",
    "alamo_el_diagonal",
    "alamo_el_fsmooth", "alamo_el_residual", "alamo_el_restrict", "alamo_el_interp_add", "alamo_el_bc_init",
    "alamo_el_newton_rhs", "alamo_el_solve", "alamo_el_norm_inf", "alamo_el_level_model", "alamo_el_coarse_rhs",
    "alamo_el_learned_input", "alamo_el_set_learned", "alamo_el_learned_guess", "alamo_amr_patch_create",
    "alamo_amr_patch_info", "alamo_amr_solve", "alamo_el_mix_models", "alamo_el_stress", "alamo_el_face_traction",
    "alamo_el_tag_bbox", "alamo_amr_fillpatch", "alamo_el_set_psi", "alamo_el_palette_size",
    "alamo_el_newton_rhs_noalloc", "alamo_el_gs_smooth", "alamo_el_tag_cells", "alamo_amr_cluster",
    "alamo_amr_average_down_coeffs", "alamo_el_level_op32",
]
LEARNED_GUESS, LEARNED_PRECOND = 1, 2

_lib = None


def lib():
    """The loaded HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AlamoError(f"{LIB_PATH} is missing: run `python -m alamo_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        L.alamo_el_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        for name in EXPORTS:
            getattr(L, name)  # raises AttributeError when a symbol is missing
        L.alamo_el_create.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(vp)]
        for name in ("alamo_el_destroy",):
            getattr(L, name).argtypes = [vp]
        L.alamo_el_set_bc.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
        L.alamo_el_set_periodic.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.alamo_el_set_uniform.argtypes = [vp, ctypes.c_int]
        L.alamo_el_set_omega.argtypes = [vp, ctypes.c_double]
        L.alamo_el_set_model.argtypes = [vp, vp, ctypes.c_int, vp]
        L.alamo_el_set_psi.argtypes = [vp, vp, ctypes.c_double, vp]
        L.alamo_el_palette_size.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.alamo_el_prepare.argtypes = [vp, vp]
        L.alamo_el_prepare_fp32.argtypes = [vp, vp]
        L.alamo_el_fapply.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.alamo_el_diagonal.argtypes = [vp, ctypes.c_int, vp, vp]
        L.alamo_el_fsmooth.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
        L.alamo_el_gs_smooth.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
        L.alamo_el_residual.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.alamo_el_restrict.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.alamo_el_interp_add.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.alamo_el_bc_init.argtypes = [vp, vp, ctypes.c_int, vp]
        L.alamo_el_newton_rhs.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.alamo_el_newton_rhs_noalloc.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.alamo_el_stress.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.alamo_el_tag_bbox.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_longlong), vp]
        L.alamo_el_face_traction.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double), vp]
        L.alamo_el_mix_models.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int, vp,
                                          vp]
        L.alamo_el_solve.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(Opts),
                                     ctypes.POINTER(Info), vp]
        L.alamo_el_norm_inf.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double), vp]
        L.alamo_el_num_levels.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.alamo_el_level_nodes.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.alamo_el_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.alamo_el_level_model.argtypes = [vp, ctypes.c_int, vp, vp]
        L.alamo_el_coarse_rhs.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.alamo_el_learned_input.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.POINTER(ctypes.c_double), vp]
        L.alamo_el_set_learned.argtypes = [vp, vp, ctypes.c_int]
        L.alamo_el_learned_guess.argtypes = [vp, vp, vp, vp]
        L.alamo_amr_patch_create.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(vp)]
        L.alamo_amr_patch_info.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 3
        L.alamo_amr_fillpatch.argtypes = [vp, vp, vp, vp, vp, vp]
        L.alamo_el_tag_cells.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, vp, vp]
        L.alamo_amr_reflux.argtypes = [vp, vp, vp, vp]
        L.alamo_amr_average_down_coeffs.argtypes = [vp, vp, vp, vp]
        L.alamo_el_level_op32.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        i3 = ctypes.POINTER(ctypes.c_int)
        L.alamo_el_box_copy.argtypes = [vp, i3, i3, vp, i3, i3, i3, i3, ctypes.c_int, vp]
        L.alamo_amr_cluster.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.alamo_amr_solve.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                      ctypes.c_double, ctypes.c_double, ctypes.POINTER(Opts), ctypes.POINTER(Info), vp]
        _lib = L
    return _lib


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise AlamoError(f"{what}: status {rc}: {lib().alamo_el_last_error().decode()}")
    return rc


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise AlamoError("no GPU: the elastic operator runs on the HIP device only")
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise AlamoError("expected a contiguous float64 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


class ElasticOperator:
    """Operator::Elastic<SYM> on one box per level (one AMR level).

    sym: ISOTROPIC, MAJORMINOR or DIAGONAL; n_cell: cells per axis; dx: cell
    size of the finest level."""

    def __init__(self, n_cell, dx, sym: int, max_coarsening_level: int = -1):
        _torch()
        n_cell = np.broadcast_to(np.asarray(n_cell, dtype=np.int32), (3,)).copy()
        dxa = np.broadcast_to(np.asarray(dx, dtype=np.float64), (3,)).copy()
        h = ctypes.c_void_p()
        _check(lib().alamo_el_create(n_cell.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     dxa.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(sym),
                                     int(max_coarsening_level), ctypes.byref(h)), "alamo_el_create")
        self._init_from(h, sym, n_cell, dxa)

    def _init_from(self, h, sym, n_cell, dxa):
        self._h = h
        self.sym = int(sym)
        self.ncomp = _model.NCOMP[self.sym]
        self.n_cell = tuple(int(v) for v in n_cell)
        self.dx = tuple(float(v) for v in dxa)
        nl = ctypes.c_int()
        _check(lib().alamo_el_num_levels(h, ctypes.byref(nl)), "alamo_el_num_levels")
        self.num_levels = nl.value
        self._bc = _bc.Constant()
        self.periodic = (0, 0, 0)
        self.parent = None
        self.origin = (0, 0, 0)       # global node index (x, y, z) of node 0
        self.dom = self.n_cell        # domain cells at this level
        self.ghost_lo = (0, 0, 0)
        self.ghost_hi = (0, 0, 0)

    def refine(self, clo, chi):
        """Patch of the next AMR level covering this level's cells [clo, chi]
        (x, y, z global indices): alamo_amr_patch_create.  The patch inherits
        the boundary table on domain faces; set its model and right-hand side
        over its own nodes (self.origin locates them)."""
        lo = np.ascontiguousarray(clo, dtype=np.int32)
        hi = np.ascontiguousarray(chi, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().alamo_amr_patch_create(self._h, lo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            hi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(h)),
               "alamo_amr_patch_create")
        p = ElasticOperator.__new__(ElasticOperator)
        o, gl, gh = (ctypes.c_int * 3)(), (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        _check(lib().alamo_amr_patch_info(h, o, gl, gh), "alamo_amr_patch_info")
        n_cell = [2 * (int(hi[a]) - int(lo[a])) + gl[a] + gh[a] for a in range(3)]
        p._init_from(h, self.sym, n_cell, np.asarray(self.dx) / 2.0)
        p.parent = self
        p.origin, p.ghost_lo, p.ghost_hi = tuple(o), tuple(gl), tuple(gh)
        p.dom = tuple(2 * d for d in self.dom)
        p.clo, p.chi = tuple(int(v) for v in lo), tuple(int(v) for v in hi)
        p._bc = self._bc
        return p

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.alamo_el_destroy(h)
            self._h = None

    # -- set-up ---------------------------------------------------------------
    def level_shape(self, mglev: int = 0):
        n = (ctypes.c_int * 3)()
        _check(lib().alamo_el_level_nodes(self._h, mglev, n), "alamo_el_level_nodes")
        return (n[2], n[1], n[0])

    def SetBC(self, bc: _bc.Constant):
        t = np.ascontiguousarray(bc.type, dtype=np.int32)
        v = np.ascontiguousarray(bc.val, dtype=np.float64)
        _check(lib().alamo_el_set_bc(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "alamo_el_set_bc")
        self._bc = bc

    def SetPeriodic(self, periodic):
        """The periodic axes (px, py, pz), the geometry's is_periodic
        (alamo_el_set_periodic): call before SetBC when the table has periodic
        entries."""
        p = np.ascontiguousarray([1 if v else 0 for v in periodic], dtype=np.int32)
        if p.shape != (3,):
            raise AlamoError("SetPeriodic: expected three flags")
        _check(lib().alamo_el_set_periodic(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
               "alamo_el_set_periodic")
        self.periodic = tuple(int(v) for v in p)

    def SetUniform(self, uniform: bool):
        _check(lib().alamo_el_set_uniform(self._h, int(bool(uniform))), "alamo_el_set_uniform")

    def SetOmega(self, omega: float):
        _check(lib().alamo_el_set_omega(self._h, float(omega)), "alamo_el_set_omega")

    def SetModel(self, C, layout: int = SOA):
        """Copy the per-node stiffness (SoA (ncomp, nz, ny, nx) or AoS
        (nz, ny, nx, ncomp)) into the operator."""
        shape = self.level_shape(0)
        want = (self.ncomp,) + shape if layout == SOA else shape + (self.ncomp,)
        if tuple(C.shape) != want:
            raise AlamoError(f"SetModel: expected shape {want}, got {tuple(C.shape)}")
        _check(lib().alamo_el_set_model(self._h, _ptr(C), int(layout), _stream()), "alamo_el_set_model")
        self._C = C

    def SetPsi(self, psi, psi_small: float = 1e-8):
        """Operator::Elastic::SetPsi (src/Operator/Elastic.H:55-57): the cell
        weight psi (nz-1, ny-1, nx-1) of the finest level."""
        shape = tuple(n - 1 for n in self.level_shape(0))
        if tuple(psi.shape) != shape:
            raise AlamoError(f"SetPsi: expected psi of shape {shape} (cells), got {tuple(psi.shape)}")
        _check(lib().alamo_el_set_psi(self._h, _ptr(psi), float(psi_small), _stream()), "alamo_el_set_psi")
        self._psi = psi

    def prepare(self):
        _check(lib().alamo_el_prepare(self._h, _stream()), "alamo_el_prepare")

    def prepare_fp32(self):
        """The FP32 copies of the FP32 cycles (alamo_el_prepare_fp32); the
        first FP32 solve after a model change makes them otherwise."""
        _check(lib().alamo_el_prepare_fp32(self._h, _stream()), "alamo_el_prepare_fp32")

    def palette_size(self, mglev: int = 0) -> int:
        """Rows of the level's C palette (0: the level reads its C copy)."""
        n = ctypes.c_int()
        _check(lib().alamo_el_palette_size(self._h, int(mglev), ctypes.byref(n)), "alamo_el_palette_size")
        return n.value

    # -- operator pieces (mglev 0 = finest) -----------------------------------
    def _v(self, t, mglev: int, ncomp: int = 3):
        """Device pointer of a field after checking its shape against the level,
        so that no kernel can be handed a short buffer."""
        if not 0 <= mglev < self.num_levels:
            raise AlamoError(f"mglev {mglev} out of range [0, {self.num_levels})")
        want = (ncomp,) + self.level_shape(mglev)
        if tuple(t.shape) != want:
            raise AlamoError(f"expected a field of shape {want}, got {tuple(t.shape)}")
        return _ptr(t)

    def _vec(self, mglev):
        torch = _torch()
        return torch.empty((3,) + self.level_shape(mglev), dtype=torch.float64, device="cuda")

    def Fapply(self, u, mglev: int = 0, out=None):
        f = self._vec(mglev) if out is None else out
        _check(lib().alamo_el_fapply(self._h, mglev, self._v(u, mglev), self._v(f, mglev), _stream()), "alamo_el_fapply")
        return f

    def Diagonal(self, mglev: int = 0):
        d = self._vec(mglev)
        _check(lib().alamo_el_diagonal(self._h, mglev, self._v(d, mglev), _stream()), "alamo_el_diagonal")
        return d

    def Fsmooth(self, x, b, mglev: int = 0, nsmooth: int = 1):
        """In place: nsmooth Fsmooth calls (2 Jacobi sweeps each)."""
        _check(lib().alamo_el_fsmooth(self._h, mglev, self._v(x, mglev), self._v(b, mglev), int(nsmooth), _stream()),
               "alamo_el_fsmooth")
        return x

    def gs_smooth(self, x, b, mglev: int = 0, nsmooth: int = 1):
        """In place: nsmooth red-black Gauss-Seidel sweeps (alamo_el_gs_smooth;
        no counterpart in the reference)."""
        _check(lib().alamo_el_gs_smooth(self._h, mglev, self._v(x, mglev), self._v(b, mglev), int(nsmooth),
                                        _stream()), "alamo_el_gs_smooth")
        return x

    def residual(self, x, b, mglev: int = 0, out=None):
        r = self._vec(mglev) if out is None else out
        _check(lib().alamo_el_residual(self._h, mglev, self._v(x, mglev), self._v(b, mglev), self._v(r, mglev), _stream()), "alamo_el_residual")
        return r

    def restriction(self, fine, cmglev: int):
        c = self._vec(cmglev)
        _check(lib().alamo_el_restrict(self._h, cmglev, self._v(c, cmglev), self._v(fine, cmglev - 1), _stream()), "alamo_el_restrict")
        return c

    def interpolation(self, fine, crse, fmglev: int):
        """In place: fine += P crse."""
        _check(lib().alamo_el_interp_add(self._h, fmglev, self._v(fine, fmglev), self._v(crse, fmglev + 1), _stream()), "alamo_el_interp_add")
        return fine

    def bc_init(self, rhs, homogeneous: bool = False):
        """BC::Operator::Elastic::Constant::Init on the finest level, in place."""
        _check(lib().alamo_el_bc_init(self._h, self._v(rhs, 0), int(bool(homogeneous)), _stream()), "alamo_el_bc_init")
        return rhs

    def newton_rhs(self, C, u, b, F0=None, layout: int = SOA, out=None, noalloc: bool = False):
        """Newton<T>::prepareForSolve right-hand side, b - N(u)
        (src/Solver/Nonlocal/Newton.H:151-231).  noalloc selects
        alamo_el_newton_rhs_noalloc (the library allocates nothing; give `out`
        so that this call does not allocate either)."""
        r = self._vec(0) if out is None else out
        f0 = self._v(F0, 0, 9) if F0 is not None else ctypes.c_void_p()
        shape = self.level_shape(0)
        want = (self.ncomp,) + shape if layout == SOA else shape + (self.ncomp,)
        if tuple(C.shape) != want:
            raise AlamoError(f"newton_rhs: expected C of shape {want}, got {tuple(C.shape)}")
        fn = lib().alamo_el_newton_rhs_noalloc if noalloc else lib().alamo_el_newton_rhs
        _check(fn(self._h, _ptr(C), int(layout), f0, self._v(u, 0), self._v(b, 0), self._v(r, 0), _stream()),
               "alamo_el_newton_rhs_noalloc" if noalloc else "alamo_el_newton_rhs")
        return r

    def mix_models(self, eta, table, out=None):
        """Integrator::Mechanics::UpdateModel on the finest level
        (src/Integrator/Mechanics.H:249-253): sum_n eta_n * table[n], from zero
        in model order.  eta: (nmodels,) + level shape, float64 on the GPU;
        table: (nmodels, ncomp) on the host (a model's C slots or its F0)."""
        import numpy as np
        tab = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
        if tab.ndim != 2:
            raise AlamoError("mix_models: table must be (nmodels, ncomp)")
        nmodels, ncomp = tab.shape
        shape = self.level_shape(0)
        if tuple(eta.shape) != (nmodels,) + shape:
            raise AlamoError(f"mix_models: expected eta of shape {(nmodels,) + shape}, got {tuple(eta.shape)}")
        torch = _torch()
        r = torch.empty((ncomp,) + shape, dtype=torch.float64, device="cuda") if out is None else out
        if tuple(r.shape) != (ncomp,) + shape:
            raise AlamoError(f"mix_models: expected out of shape {(ncomp,) + shape}, got {tuple(r.shape)}")
        _check(lib().alamo_el_mix_models(self._h, _ptr(eta), int(nmodels),
                                         tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(ncomp), _ptr(r),
                                         _stream()), "alamo_el_mix_models")
        return r

    def stress(self, C, u, F0=None, layout: int = SOA, strain: bool = False):
        """Base::Mechanics stress (and strain) on the finest level
        (src/Integrator/Base/Mechanics.H:210-245): sigma = C:(grad u - F0),
        9 planes (3p+q).  Returns sigma, or (sigma, strain)."""
        torch = _torch()
        shape = self.level_shape(0)
        want = (self.ncomp,) + shape if layout == SOA else shape + (self.ncomp,)
        if tuple(C.shape) != want:
            raise AlamoError(f"stress: expected C of shape {want}, got {tuple(C.shape)}")
        f0 = self._v(F0, 0, 9) if F0 is not None else ctypes.c_void_p()
        sig = torch.empty((9,) + shape, dtype=torch.float64, device="cuda")
        eps = torch.empty((9,) + shape, dtype=torch.float64, device="cuda") if strain else None
        _check(lib().alamo_el_stress(self._h, _ptr(C), int(layout), f0, self._v(u, 0), _ptr(sig),
                                     _ptr(eps) if strain else ctypes.c_void_p(), _stream()), "alamo_el_stress")
        return (sig, eps) if strain else sig

    def face_traction(self, sigma, u, axis: int = 0, side: int = 1):
        """Base::Mechanics::Integrate, 3D (src/Integrator/Base/Mechanics.H:391-448):
        (traction, displacement) on a domain face; axis 0, side 1 is the
        reference's trac_hi / disp_hi of thermo.dat."""
        import numpy as np
        trac = (ctypes.c_double * 3)()
        disp = (ctypes.c_double * 3)()
        _check(lib().alamo_el_face_traction(self._h, self._v(sigma, 0, 9), self._v(u, 0), int(axis), int(side), trac,
                                            disp, _stream()), "alamo_el_face_traction")
        return np.array(trac[:]), np.array(disp[:])

    def tag_box(self, eta, threshold: float = 0.1, buffer: int = 2):
        """Integrator::Mechanics::TagCellsForRefinement on the finest level
        (src/Integrator/Mechanics.H:298-313): ((clo, chi), ntag) with [clo, chi]
        the tagged cells' bounding box grown by `buffer` cells (level cell
        indices), or (None, 0) when no cell is tagged.  eta: (ncomp,) + level
        shape, float64 on the GPU.  threshold is eta_ref_threshold (0.01 by
        default in the reference, src/Integrator/Mechanics.H:137; 0.1 for the
        three-inclusion AMR case)."""
        shape = self.level_shape(0)
        if eta.dim() != 4 or tuple(eta.shape[1:]) != shape:
            raise AlamoError(f"tag_box: expected eta of shape (ncomp,) + {shape}, got {tuple(eta.shape)}")
        clo = (ctypes.c_int * 3)()
        chi = (ctypes.c_int * 3)()
        n = ctypes.c_longlong()
        _check(lib().alamo_el_tag_bbox(self._h, _ptr(eta), int(eta.shape[0]), float(threshold), int(buffer), clo,
                                       chi, ctypes.byref(n), _stream()), "alamo_el_tag_bbox")
        if n.value == 0:
            return None, 0
        return (tuple(clo[:]), tuple(chi[:])), int(n.value)

    def child_box(self, eta, threshold: float = 0.1, buffer: int = 2, min_cells: int = 2,
                  blocking_factor: int = 0):
        """Cells [clo, chi] (global, this level) of the next level's patch from
        this level's tags: tag_box's bounding box, widened so that the patch's
        extent on the next level is a multiple of blocking_factor (amr.blocking_factor,
        when above 1) and clipped so that the patch and its ghost layer lie on
        this handle's unknowns (the nesting alamo_amr_patch_create requires);
        None when nothing is tagged or fewer than min_cells cells remain along
        an axis."""
        box, _ = self.tag_box(eta, threshold, buffer)
        if box is None:
            return None
        nodes = self.level_shape(0)[::-1]
        q = max(1, int(blocking_factor) // 2)
        clo, chi = [], []
        for a in range(3):
            ulo = self.origin[a] + self.ghost_lo[a]
            uhi = self.origin[a] + nodes[a] - 1 - self.ghost_hi[a]
            blo = (box[0][a] + self.origin[a]) // q * q
            bhi = -(-(box[1][a] + self.origin[a]) // q) * q
            lo = max(blo, ulo + (1 if ulo > 0 else 0))
            hi = min(bhi, uhi - (1 if uhi < self.dom[a] else 0))
            if hi - lo < min_cells:
                return None
            clo.append(lo)
            chi.append(hi)
        return tuple(clo), tuple(chi)

    def tag_cells(self, eta, threshold: float = 0.1):
        """The cell tags of tag_box before the bounding box: a uint8 tensor
        (nz-1, ny-1, nx-1) on the GPU, 1 where a cell is tagged
        (alamo_el_tag_cells)."""
        shape = self.level_shape(0)
        if eta.dim() != 4 or tuple(eta.shape[1:]) != shape:
            raise AlamoError(f"tag_cells: expected eta of shape (ncomp,) + {shape}, got {tuple(eta.shape)}")
        torch = _torch()
        tags = torch.empty(tuple(n - 1 for n in shape), dtype=torch.uint8, device=eta.device)
        _check(lib().alamo_el_tag_cells(self._h, _ptr(eta), int(eta.shape[0]), float(threshold),
                                        ctypes.c_void_p(tags.data_ptr()), _stream()), "alamo_el_tag_cells")
        return tags

    def child_boxes(self, eta, threshold: float = 0.1, buffer: int = 2, blocking_factor: int = 0,
                    grid_eff: float = 0.7, min_cells: int = 2):
        """Cells [(clo, chi)] (global, this level) of the next level's patches
        from this level's tags, clustered into several boxes
        (alamo_amr_cluster: Berger-Rigoutsos with grid_eff, boxes aligned to
        the blocking factor, boxes closer than one cell merged), each clipped
        so that the patch and its ghost layer lie on this handle's unknowns;
        boxes thinner than min_cells along an axis are dropped.  Only cells
        among this patch's own unknowns (not its ghost layer) are tagged."""
        import numpy as np
        tags = self.tag_cells(eta, threshold).cpu().numpy()
        nodes = self.level_shape(0)[::-1]
        # cells of the handle that lie among this patch's own cells
        lo = [self.ghost_lo[a] for a in range(3)]
        hi = [nodes[a] - 1 - self.ghost_hi[a] for a in range(3)]  # exclusive cell bound
        sub = np.ascontiguousarray(tags[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
        if sub.size == 0 or not sub.any():
            return []
        nc = (ctypes.c_int * 3)(sub.shape[2], sub.shape[1], sub.shape[0])
        org = (ctypes.c_int * 3)(*[self.origin[a] + lo[a] for a in range(3)])
        cap = 4096
        out = (ctypes.c_int * (6 * cap))()
        nb = ctypes.c_int()
        _check(lib().alamo_amr_cluster(sub.ctypes.data, nc, org, int(buffer), int(blocking_factor), float(grid_eff),
                                       cap, out, ctypes.byref(nb)), "alamo_amr_cluster")
        boxes = []
        for q in range(nb.value):
            clo, chi = [], []
            for a in range(3):
                ulo = self.origin[a] + self.ghost_lo[a]
                uhi = self.origin[a] + nodes[a] - 1 - self.ghost_hi[a]
                clo.append(max(out[6 * q + a], ulo + (1 if ulo > 0 else 0)))
                chi.append(min(out[6 * q + 3 + a], uhi - (1 if uhi < self.dom[a] else 0)))
            if all(chi[a] - clo[a] >= min_cells for a in range(3)):
                boxes.append((tuple(clo), tuple(chi)))
        return boxes

    def reflux(self, fine_res, parent_res):
        """Operator::reflux of this patch into its parent (alamo_amr_reflux):
        parent_res (a field of the parent) at the covered and coarse-fine
        nodes becomes the full weighting of fine_res (ghost rows read as 0)."""
        if self.parent is None:
            raise AlamoError("reflux: this operator is not a patch")
        self._check_field(fine_res, self.level_shape(0), "fine_res")
        self._check_field(parent_res, self.parent.level_shape(0), "parent_res")
        _check(lib().alamo_amr_reflux(self._h, _ptr(fine_res), _ptr(parent_res), _stream()), "alamo_amr_reflux")
        return parent_res

    def average_down_coeffs(self, C, parent_C):
        """Elastic::averageDownCoeffsDifferentAmrLevels (src/Operator/
        Elastic.cpp:628-745) of this patch (alamo_amr_average_down_coeffs):
        parent_C (SoA over the parent's nodes) at the covered and coarse-fine
        nodes becomes the full weighting of C (SoA over this patch's nodes,
        ghost layer included).  Set the parent's model from parent_C after."""
        if self.parent is None:
            raise AlamoError("average_down_coeffs: this operator is not a patch")
        for v, shape, name in ((C, self.level_shape(0), "C"), (parent_C, self.parent.level_shape(0), "parent_C")):
            if tuple(v.shape) != (self.ncomp,) + tuple(shape) or v.dtype != _torch().float64 or not v.is_contiguous():
                raise AlamoError(f"average_down_coeffs: expected {name} as a contiguous float64 tensor of shape "
                                 f"{(self.ncomp,) + tuple(shape)}, got {tuple(v.shape)} {v.dtype}")
        _check(lib().alamo_amr_average_down_coeffs(self._h, _ptr(C), _ptr(parent_C), _stream()),
               "alamo_amr_average_down_coeffs")
        return parent_C

    def fill_patch(self, parent_u, old=None, old_u=None):
        """This patch's field after a regrid (alamo_amr_fillpatch;
        Integrator::FillPatch / FillCoarsePatch, src/Integrator/Integrator.cpp:
        394-478): the old patch's values where its unknowns hold a node, else
        the parent field parent_u interpolated trilinearly.  Returns a new
        (3,) + level shape float64 tensor on the GPU."""
        if self.parent is None:
            raise AlamoError("fill_patch: this operator is not a patch")
        if (old is None) != (old_u is None):
            raise AlamoError("fill_patch: give both the old patch and its field, or neither")
        shape = self.level_shape(0)
        self._check_field(parent_u, self.parent.level_shape(0), "parent_u")
        if old is not None:
            self._check_field(old_u, old.level_shape(0), "old_u")
        out = _torch().empty((3,) + shape, dtype=_torch().float64, device=parent_u.device)
        _check(lib().alamo_amr_fillpatch(self._h, _ptr(parent_u), old._h if old is not None else None,
                                         _ptr(old_u) if old is not None else None, _ptr(out), _stream()),
               "alamo_amr_fillpatch")
        return out

    @staticmethod
    def _check_field(v, shape, name):
        if tuple(v.shape) != (3,) + tuple(shape) or v.dtype != _torch().float64 or not v.is_contiguous():
            raise AlamoError(f"{name}: expected a contiguous float64 tensor of shape {(3,) + tuple(shape)}, "
                             f"got {tuple(v.shape)} {v.dtype}")

    OP32 = {"fapply": 0, "residual": 1, "jacobi": 2, "restrict": 3, "interp_add": 4}

    def level_op32(self, op: str, mglev: int, x, b=None, out=None):
        """One FP32 kernel of the FP32 cycles on MG level mglev
        (alamo_el_level_op32): "fapply", "residual" (b - A x), "jacobi" (one
        sweep), "restrict" (x onto mglev + 1), "interp_add" (out += P x, x on
        mglev + 1).  Float32 CUDA tensors (3, nz, ny, nx)."""
        torch = _torch()
        code = self.OP32[op]
        src_lev = mglev + 1 if code == 4 else mglev
        dst_lev = mglev + 1 if code == 3 else mglev
        for v, lev, name in ((x, src_lev, "x"), (b, mglev, "b"), (out, dst_lev, "out")):
            if v is None:
                continue
            want = (3,) + tuple(self.level_shape(lev))
            if tuple(v.shape) != want or v.dtype != torch.float32 or not v.is_contiguous():
                raise AlamoError(f"level_op32: expected {name} as a contiguous float32 tensor of shape {want}")
        if out is None:
            out = torch.zeros((3,) + tuple(self.level_shape(dst_lev)), dtype=torch.float32, device="cuda")
        def p32(t):
            return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()

        if not x.is_cuda or (b is not None and not b.is_cuda) or not out.is_cuda:
            raise AlamoError("level_op32: expected CUDA tensors")
        _check(lib().alamo_el_level_op32(self._h, int(mglev), code, p32(x), p32(b), p32(out), _stream()),
               "alamo_el_level_op32")
        return out

    def norm_inf(self, v, mglev: int = 0) -> float:
        out = ctypes.c_double()
        _check(lib().alamo_el_norm_inf(self._h, mglev, self._v(v, mglev), ctypes.byref(out), _stream()), "alamo_el_norm_inf")
        return out.value

    # --- learned coarse-level solver (no counterpart in the reference) -----
    def level_model(self, mglev: int):
        """Stiffness of MG level mglev, SoA (ncomp, nz, ny, nx)."""
        torch = _torch()
        C = torch.empty((self.ncomp,) + self.level_shape(mglev), dtype=torch.float64, device="cuda")
        _check(lib().alamo_el_level_model(self._h, mglev, _ptr(C), _stream()), "alamo_el_level_model")
        return C

    def coarse_rhs(self, b, mglev: int):
        """Level-mglev right-hand side of the finest-level b."""
        torch = _torch()
        bc = torch.empty((3,) + self.level_shape(mglev), dtype=torch.float64, device="cuda")
        _check(lib().alamo_el_coarse_rhs(self._h, mglev, self._v(b, 0), self._v(bc, mglev), _stream()),
               "alamo_el_coarse_rhs")
        return bc

    def learned_input(self, bc, mglev: int):
        """(features float32 (ncomp + 30, nz, ny, nx), scale s) of the
        level-mglev right-hand side bc; the network predicts u / s."""
        torch = _torch()
        f = torch.empty((self.ncomp + 30,) + self.level_shape(mglev), dtype=torch.float32, device="cuda")
        sc = ctypes.c_double()
        _check(lib().alamo_el_learned_input(self._h, mglev, self._v(bc, mglev), ctypes.c_void_p(f.data_ptr()),
                                            ctypes.byref(sc), _stream()), "alamo_el_learned_input")
        return f, sc.value

    def set_learned(self, net, mglev: int):
        """Attach a learned.UNet working on MG level mglev (None detaches)."""
        _check(lib().alamo_el_set_learned(self._h, net.h if net is not None else None, int(mglev)),
               "alamo_el_set_learned")
        self._net = net  # the handle keeps a pointer: keep the network alive

    def learned_guess(self, b):
        torch = _torch()
        x = torch.empty_like(b)
        _check(lib().alamo_el_learned_guess(self._h, self._v(b, 0), self._v(x, 0), _stream()),
               "alamo_el_learned_guess")
        return x


def default_opts() -> Opts:
    o = Opts()
    lib().alamo_el_default_opts(ctypes.byref(o))
    return o


# The MI355X fast path (FP32 cycles inside FP64 FGMRES): Linear(op, **FAST_PATH),
# or Linear(op, **fast_path(nodes)) for the size rule of max_fmg_iter below.
# FGMRES(10) (the Gram-Schmidt passes read every basis vector, so a short basis
# is cheaper: 10 took the steps of 30 in less time at 129^3); the
# preconditioner runs 3 + 3 Fsmooth calls (16 steps at 129^3 against 23 with
# the reference's 2 + 2); the direct FP64 bottom.
FAST_PATH = dict(cycle="F", precision="fp32", outer="fgmres", restart=10, bottom_solver="direct", pre_smooth=3,
                 post_smooth=3)


def fast_path(nodes: int) -> dict:
    """FAST_PATH with the F-cycle schedule for a finest level of `nodes` nodes.
    An F-cycle's extra visits of the coarse levels cost a fixed, latency-bound
    time, while its better convergence saves steps whose cost grows with the
    grid: below 8 M nodes F-cycles for the first two steps and V-cycles after
    are fastest (129^3: 28.2 ms in 22 steps against 30.6 ms in 16 with
    F-cycles throughout), above it F-cycles throughout (257^3: 0.166-0.184 s
    in 17-18 steps against 0.20 s in 26 with the switch after two steps and
    0.18 s after ten)."""
    return dict(FAST_PATH, max_fmg_iter=2 if nodes < 8_000_000 else -1)


class Linear:
    """Solver::Nonlocal::Linear: the MLMG solve with the reference defaults
    (src/Solver/Nonlocal/Linear.H:255-290).  Keyword names follow the
    reference's solver.* inputs (Linear::Parse, Linear.H:296-382) plus
    ``cycle`` ("F" = setMaxFmgIter(1e8), "V"), ``precision`` ("fp64",
    "fp32"), ``outer`` ("mlmg", "fgmres") and ``bottom_solver`` ("bicgstab",
    the reference's; "direct": the cycles stop at the finest MG level with at
    most 4096 unknowns and solve there with the FP64 inverse of its operator) and
    ``learned`` ("", "guess", "guess+precond": the learned coarse-level solver
    attached with ElasticOperator.set_learned) and ``smoother`` ("jacobi", the
    reference's damped Jacobi; "rbgs": red-black Gauss-Seidel)."""

    def __init__(self, op: ElasticOperator, max_iter=200, pre_smooth=2, post_smooth=2, bottom_max_iter=200,
                 final_smooth=8, bottom_tol_rel=1e-4, omega=2.0 / 3.0, cycle="F", precision="fp64", outer="mlmg",
                 restart=30, verbose=0, abort_on_fail=True, bottom_solver="bicgstab", learned="", smoother="jacobi",
                 max_fmg_iter=-1, fixed_iter=-1):
        self.op = op
        o = default_opts()
        o.max_iter, o.nu1, o.nu2 = int(max_iter), int(pre_smooth), int(post_smooth)
        o.bottom_max_iter, o.final_smooth = int(bottom_max_iter), int(final_smooth)
        o.bottom_tol_rel, o.omega = float(bottom_tol_rel), float(omega)
        o.cycle = {"F": CYCLE_F, "V": CYCLE_V}[cycle]
        o.precision = {"fp64": PREC_FP64, "fp32": PREC_FP32_IR}[precision]
        o.outer = {"mlmg": OUTER_MLMG, "fgmres": OUTER_FGMRES}[outer]
        o.restart, o.verbose = int(restart), int(verbose)
        o.bottom_solver = {"bicgstab": 0, "direct": 1}[bottom_solver]
        o.smoother = {"jacobi": 0, "rbgs": 1}[smoother]
        o.max_fmg_iter, o.fixed_iter = int(max_fmg_iter), int(fixed_iter)
        o.learned = {"": 0, "guess": LEARNED_GUESS, "precond": LEARNED_PRECOND,
                     "guess+precond": LEARNED_GUESS | LEARNED_PRECOND}[learned]
        self.opts = o
        self.abort_on_fail = abort_on_fail
        self.info = Info()

    def solve(self, x, b, tol_rel: float, tol_abs: float) -> float:
        """Solve A x = b in place; returns the final residual inf-norm
        (the return value of Linear::solve)."""
        rc = lib().alamo_el_solve(self.op._h, self.op._v(x, 0), self.op._v(b, 0), float(tol_rel), float(tol_abs),
                                  ctypes.byref(self.opts), ctypes.byref(self.info), _stream())
        _check(rc, "alamo_el_solve")
        if rc == NOT_CONVERGED and self.abort_on_fail:
            raise SolverError(f"MLMG failed to converge: |r| = {self.info.res:.3e} after {self.info.iters} iterations")
        return self.info.res


def amr_levels(base, eta_of, levels: int = 3, threshold: float = 0.1, buffer: int = 2, blocking_factor: int = 0,
               grid_eff=None):
    """AMR levels from the device tags (Integrator::Mechanics::
    TagCellsForRefinement, src/Integrator/Mechanics.H:298-313): starting from
    the base operator, each level tags its cells from eta_of(op), the
    (ncomp,) + level shape indicator field on op's nodes (float64, GPU), up to
    `levels` levels.  With grid_eff None every level is one patch, the clipped
    bounding box of the tags (child_box); with a grid_eff every patch's tags are
    clustered into several boxes (child_boxes).  Returns [base, patches...],
    parents before children; set each patch's model."""
    ops = [base]
    front = [base]
    for _ in range(levels - 1):
        nxt = []
        for op in front:
            if grid_eff is None:
                box = op.child_box(eta_of(op), threshold, buffer, blocking_factor=blocking_factor)
                boxes = [] if box is None else [box]
            else:
                boxes = op.child_boxes(eta_of(op), threshold, buffer, blocking_factor=blocking_factor,
                                       grid_eff=grid_eff)
            nxt += [op.refine(*b) for b in boxes]
        if not nxt:
            break
        ops += nxt
        front = nxt
    return ops


def _depth(op):
    return 0 if op.parent is None else 1 + _depth(op.parent)


def amr_regrid(old_ops, old_xs, eta_of, levels: int = 3, threshold: float = 0.1, buffer: int = 2,
               blocking_factor: int = 0, grid_eff=None):
    """Regrid (AmrCore::regrid with Integrator::RemakeLevel / MakeNewLevelFromCoarse,
    src/Integrator/Integrator.cpp:394-478): new patches over the same base from
    the tags (amr_levels), and each new patch's field filled, parents first:
    the already regridded parent's field interpolated, overwritten by the
    values of every old patch of the same level where that patch's unknowns
    reach (fill_patch).  Returns (ops, xs) with xs[0] = old_xs[0]."""
    ops = amr_levels(old_ops[0], eta_of, levels, threshold, buffer, blocking_factor, grid_eff)
    xs = [old_xs[0]]
    for op in ops[1:]:
        parent_u = xs[next(q for q, o in enumerate(ops) if o is op.parent)]
        u = op.fill_patch(parent_u)
        nodes = op.level_shape(0)[::-1]
        d = _depth(op)
        for old, old_u in zip(old_ops[1:], old_xs[1:]):
            if _depth(old) != d:
                continue
            lo = [max(2 * old.clo[a], op.origin[a]) - op.origin[a] for a in range(3)]
            hi = [min(2 * old.chi[a], op.origin[a] + nodes[a] - 1) - op.origin[a] for a in range(3)]
            if any(hi[a] < lo[a] for a in range(3)):
                continue
            t = op.fill_patch(parent_u, old, old_u)
            sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            u[sl] = t[sl]
        xs.append(u)
    return ops, xs


def amr_solve(levels, xs, bs, tol_rel: float, tol_abs: float, **kw):
    """Composite solve over a base operator and its nested patches
    (alamo_amr_solve); xs, bs are per-level fields, solved in place.  Keyword
    options as for Linear.  Returns the Info structure."""
    if not (len(levels) == len(xs) == len(bs)):
        raise AlamoError("levels, xs and bs must have the same length")
    solver = Linear(levels[0], **kw)
    n = len(levels)
    H = (ctypes.c_void_p * n)(*[lv._h.value for lv in levels])
    X = (ctypes.c_void_p * n)(*[lv._v(x, 0).value for lv, x in zip(levels, xs)])
    B = (ctypes.c_void_p * n)(*[lv._v(b, 0).value for lv, b in zip(levels, bs)])
    rc = lib().alamo_amr_solve(n, H, X, B, float(tol_rel), float(tol_abs), ctypes.byref(solver.opts),
                               ctypes.byref(solver.info), _stream())
    _check(rc, "alamo_amr_solve")
    if rc == NOT_CONVERGED and solver.abort_on_fail:
        raise SolverError(f"composite MLMG failed to converge: |r| = {solver.info.res:.3e}")
    return solver.info


class Newton(Linear):
    """Newton<T>::solve for linear models (nriters = 1, Newton.H:630): builds
    rhs = b - N(u) with DW = C:(grad u - F0), solves A du = rhs, u += du."""

    def solve(self, u, b, C, tol_rel: float, tol_abs: float, F0=None, layout: int = SOA) -> float:  # noqa: D401
        torch = _torch()
        rhs = self.op.newton_rhs(C, u, b, F0=F0, layout=layout)
        du = torch.zeros_like(u)
        res = Linear.solve(self, du, rhs, tol_rel, tol_abs)
        u += du
        return res


def amr_average_down_coeffs(ops, Cs):
    """Linear's average_down_coeffs over a hierarchy (Elastic::averageDownCoeffs,
    src/Operator/Elastic.cpp:605-614): every patch's C restricted onto its
    parent's, the deepest patches first, then every model set again.  ops as
    amr_levels returns them (parents before children); Cs their SoA C tensors,
    changed in place."""
    idx = {id(op): q for q, op in enumerate(ops)}
    for q in sorted(range(1, len(ops)), key=lambda q: -_depth(ops[q])):
        ops[q].average_down_coeffs(Cs[q], Cs[idx[id(ops[q].parent)]])
    for op, C in zip(ops, Cs):
        op.SetModel(C)
    return Cs


def box_copy(dst, dst_origin, src, src_origin, lo, hi):
    """alamo_el_box_copy: dst[:, nodes lo..hi] = src[:, same nodes], the
    arrays (ncomp, nz, ny, nx) float64 on the GPU with their node 0 at the
    global (x, y, z) origins given; lo, hi global node indices (x, y, z),
    inclusive.  Gathers the FABs of a multi-box field into a level's single box,
    or scatters them back."""
    if dst.shape[0] != src.shape[0]:
        raise AlamoError("box_copy: the arrays must hold the same number of components")
    for t in (dst, src):
        if t.dim() != 4 or t.dtype != _torch().float64 or not t.is_contiguous() or not t.is_cuda:
            raise AlamoError("box_copy: expected contiguous float64 CUDA tensors (ncomp, nz, ny, nx)")
    c3 = lambda v: (ctypes.c_int * 3)(*[int(a) for a in v])
    _check(lib().alamo_el_box_copy(ctypes.c_void_p(dst.data_ptr()), c3(dst.shape[1:][::-1]), c3(dst_origin),
                                   ctypes.c_void_p(src.data_ptr()), c3(src.shape[1:][::-1]), c3(src_origin), c3(lo),
                                   c3(hi), int(dst.shape[0]), _stream()), "alamo_el_box_copy")
    return dst
