"""ctypes binding of libhyteg_amd.so (the C ABI declared in include/hyteg_amd.h).

The binding is plumbing for tests and benchmarks: every operation runs in the HIP library.
Loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libhyteg_amd.so")
HEADER = os.path.join(ROOT, "include", "hyteg_amd.h")
MESHES = os.path.join(ROOT, "meshes")

OK, ERR_ARG, ERR_HIP, ERR_IO, ERR_NUMERIC, ERR_STATE = 0, 1, 2, 3, 4, 5
NONE, INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 0, 1, 2, 4, 8, 15
REPLACE, ADD = 0, 1
F64, F32 = 8, 4

_lib = None


class HygError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hyg error {code}: {msg}")
        self.code = code


def _sig(lib):
    vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    P = np.ctypeslib.ndpointer
    pvp = C.POINTER(C.c_void_p)
    table = {
        "hyg_last_error": (C.c_char_p, []),
        "hyg_init": (i, [i]),
        "hyg_mesh_create_from_msh": (i, [C.c_char_p, pvp]),
        "hyg_mesh_create_from_msh_tags": (i, [C.c_char_p, i, pvp]),
        "hyg_mesh_create_spherical_shell": (i, [i, i, d, d, pvp]),
        "hyg_mesh_create_symmetric_cuboid": (i, [C.POINTER(C.c_double), C.POINTER(C.c_double), i, i, i, pvp]),
        "hyg_mesh_create_cuboid": (i, [C.POINTER(C.c_double), C.POINTER(C.c_double), i, i, i, pvp]),
        "hyg_mesh_create_cubed_domain": (i, [P(np.int32), i, i, pvp]),
        "hyg_mesh_destroy": (i, [vp]),
        "hyg_mesh_info": (i, [vp, P(np.int64)]),
        "hyg_mesh_num_dofs": (i, [vp, i, C.POINTER(C.c_int64)]),
        "hyg_mesh_dof_coords": (i, [vp, i, vp, vp]),
        "hyg_func_create": (i, [vp, i, i, i, pvp]),
        "hyg_func_create_p2": (i, [vp, i, i, i, pvp]),
        "hyg_func_destroy": (i, [vp]),
        "hyg_func_upload": (i, [vp, i, P(np.float64), vp]),
        "hyg_func_download": (i, [vp, i, P(np.float64), vp]),
        "hyg_interpolate_const": (i, [vp, d, i, i, vp]),
        "hyg_assign": (i, [vp, i, P(np.float64), pvp, i, i, vp]),
        "hyg_add": (i, [vp, i, P(np.float64), pvp, i, i, vp]),
        "hyg_mult_elementwise": (i, [vp, vp, vp, i, i, vp]),
        "hyg_dot": (i, [vp, vp, i, i, C.POINTER(C.c_double), vp]),
        "hyg_dot_noalloc": (i, [vp, vp, i, i, vp, vp]),
        "hyg_halo_layout": (i, [vp, i, P(np.int64)]),
        "hyg_halo_pack": (i, [vp, i, vp, vp]),
        "hyg_halo_unpack": (i, [vp, i, vp, vp]),
        "hyg_halo_sum_replicas": (i, [vp, i, i, vp]),
        "hyg_batch_create": (i, [i, pvp, pvp, pvp, pvp]),
        "hyg_batch_destroy": (i, [vp]),
        "hyg_batch_set_config": (i, [vp, i]),
        "hyg_batch_vcycle": (i, [vp, i, vp]),
        "hyg_batch_apply": (i, [vp, pvp, pvp, i, vp]),
        "hyg_batch_residual_norms": (i, [vp, P(np.float64), vp]),
        "hyg_batch_ir_solve": (i, [vp, pvp, pvp, i, i, d, i, P(np.float64), C.POINTER(C.c_int), vp]),
        "hyg_convert": (i, [vp, vp, i, vp]),
        "hyg_op_create_p1_const": (i, [vp, i, i, pvp]),
        "hyg_op_create_p1_divkgrad": (i, [vp, i, i, d, P(np.float64), i, pvp]),
        "hyg_op_create_p1_divkgrad_poly": (i, [vp, i, i, i, P(np.float64), pvp]),
        "hyg_op_interpolate_stencils": (i, [vp, i, i]),
        "hyg_mesh_set_shell_map": (i, [vp]),
        "hyg_op_create_p1_blending_laplace": (i, [vp, i, i, pvp]),
        "hyg_op_create_p2_const": (i, [vp, i, i, pvp]),
        "hyg_op_create_p2_divkgrad": (i, [vp, i, i, i, d, d, d, C.c_uint64, pvp]),
        "hyg_op_store_local_matrices": (i, [vp]),
        "hyg_apply_vector": (i, [vp, i, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), i, i, i, vp]),
        "hyg_op_destroy": (i, [vp]),
        "hyg_op_cell_stencil": (i, [vp, i, i, P(np.float64)]),
        "hyg_apply": (i, [vp, vp, vp, i, i, i, vp]),
        "hyg_op_tune": (i, [vp, i, i, i, vp, vp]),
        "hyg_op_set_config": (i, [vp, i, i, i, i]),
        "hyg_op_get_config": (i, [vp, i, i, i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "hyg_smooth_jac": (i, [vp, vp, vp, vp, d, i, i, vp]),
        "hyg_smooth_sor": (i, [vp, vp, vp, d, i, i, i, vp]),
        "hyg_smooth_sor_ordered": (i, [vp, vp, vp, d, i, i, i, vp]),
        "hyg_mesh_sor_colours": (i, [vp, C.POINTER(C.c_int), P(np.int32)]),
        "hyg_inverse_diagonal": (i, [vp, vp, i, vp]),
        "hyg_residual": (i, [vp, vp, vp, vp, i, i, vp]),
        "hyg_restrict": (i, [vp, i, i, vp]),
        "hyg_prolongate": (i, [vp, i, i, i, vp]),
        "hyg_cheb_coefficients": (i, [i, d, d, P(np.float64)]),
        "hyg_power_iteration": (i, [vp, vp, vp, i, i, C.POINTER(C.c_double), vp]),
        "hyg_gmg_create": (i, [vp, i, P(np.int32), P(np.float64), pvp]),
        "hyg_gmg_destroy": (i, [vp]),
        "hyg_gmg_vcycle": (i, [vp, vp, vp, i, vp]),
        "hyg_gmg_vcycle_graph": (i, [vp, vp, vp, i, i, vp]),
        "hyg_gmg_solve": (i, [vp, vp, vp, i, i, P(np.float64), vp]),
        "hyg_gmg_solve_tol": (i, [vp, vp, vp, i, i, d, P(np.float64), C.POINTER(C.c_int), vp]),
        "hyg_ir_solve": (i, [vp, vp, vp, vp, i, i, i, d, P(np.float64), C.POINTER(C.c_int), vp]),
    }
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return table


def header_symbols():
    """Entry points declared in include/hyteg_amd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hyg_\w+)\s*\(", text, re.M)))


def load():
    """Load the HIP library (no device calls)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        _sig(lib)
        _lib = lib
    return _lib


def check(rc):
    if rc != OK:
        raise HygError(rc, load().hyg_last_error().decode())


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(load().hyg_init(device))
        _initialised = True


def _out():
    return C.c_void_p()


class Mesh:
    """A Gmsh v2.2 file (name or path), the spherical shell ("shell", ntan, nrad, rmin, rmax), the symmetric
    cuboid ("symmetric_cuboid", lower xyz, upper xyz, nx, ny, nz) the six-tet cuboid ("cuboid", same arguments) or a cubed domain ("cubed_domain", integer cube positions,
    cube type 0 or 1)."""

    def __init__(self, name, *args, tags=False):
        """tags=True (Gmsh files): the file's physical tags are the mesh boundary flags (create0123BC mapping)."""
        init()
        h = _out()
        if name == "shell":
            self.path = None
            ntan, nrad, rmin, rmax = args
            check(load().hyg_mesh_create_spherical_shell(ntan, nrad, rmin, rmax, C.byref(h)))
        elif name in ("symmetric_cuboid", "cuboid"):
            self.path = None
            lo, hi, nx, ny, nz = args
            lo3, hi3 = (C.c_double * 3)(*lo), (C.c_double * 3)(*hi)
            fn = load().hyg_mesh_create_symmetric_cuboid if name == "symmetric_cuboid" else load().hyg_mesh_create_cuboid
            check(fn(lo3, hi3, nx, ny, nz, C.byref(h)))
        elif name == "cubed_domain":
            self.path = None
            cubes, cube_type = args
            cb = np.ascontiguousarray(np.asarray(cubes, dtype=np.int32).reshape(-1))
            check(load().hyg_mesh_create_cubed_domain(cb, len(cb) // 3, cube_type, C.byref(h)))
        else:
            self.path = name if os.path.isabs(name) else os.path.join(MESHES, name)
            check(load().hyg_mesh_create_from_msh_tags(self.path.encode(), 1 if tags else 0, C.byref(h)))
        self.h = h

    def info(self):
        out = np.zeros(7, dtype=np.int64)
        check(load().hyg_mesh_info(self.h, out))
        return out

    def ndofs(self, L):
        n = C.c_int64()
        check(load().hyg_mesh_num_dofs(self.h, L, C.byref(n)))
        return n.value

    def coords(self, L):
        n = self.ndofs(L)
        xyz = np.zeros(3 * n)
        bnd = np.zeros(n, dtype=np.uint8)
        check(load().hyg_mesh_dof_coords(self.h, L, xyz.ctypes.data, bnd.ctypes.data))
        return xyz.reshape(n, 3), bnd.astype(bool)

    def set_shell_map(self):
        """IcosahedralShellMap::setMap: blended cells (coords() returns blended positions from then on)."""
        check(load().hyg_mesh_set_shell_map(self.h))

    def sor_colours(self):
        """(number of colours, colour of every macro-vertex) of the multi-colour SOR ordering."""
        n = C.c_int()
        nv = int(self.info()[0])
        cols = np.zeros(max(nv, 1), dtype=np.int32)
        check(load().hyg_mesh_sor_colours(self.h, C.byref(n), cols))
        return n.value, cols[:nv]

    def boundary(self, L):
        bnd = np.zeros(self.ndofs(L), dtype=np.uint8)
        check(load().hyg_mesh_dof_coords(self.h, L, None, bnd.ctypes.data))
        return bnd.astype(bool)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hyg_mesh_destroy(self.h)


class Function:
    """P1 function (space="p1") or P2 function (space="p2"); a P2 function of level L exchanges the values
    of the vertex lattice of level L+1 (mesh.ndofs(L + 1) values, canonical order of mesh.coords(L + 1))."""

    def __init__(self, mesh, minL, maxL, dtype=F64, space="p1"):
        self.mesh, self.minL, self.maxL, self.dtype = mesh, minL, maxL, dtype
        self.lat = 1 if space == "p2" else 0
        h = _out()
        create = load().hyg_func_create_p2 if self.lat else load().hyg_func_create
        check(create(mesh.h, minL, maxL, dtype, C.byref(h)))
        self.h = h

    def upload(self, L, v):
        check(load().hyg_func_upload(self.h, L, np.ascontiguousarray(v, dtype=np.float64), None))

    def download(self, L):
        out = np.zeros(self.mesh.ndofs(L + self.lat))
        check(load().hyg_func_download(self.h, L, out, None))
        return out

    def interpolate(self, value, L, flag=ALL):
        check(load().hyg_interpolate_const(self.h, value, L, flag, None))

    def assign(self, scalars, funcs, L, flag=ALL):
        arr = (C.c_void_p * len(funcs))(*[f.h.value for f in funcs])
        check(load().hyg_assign(self.h, len(funcs), np.asarray(scalars, dtype=np.float64), arr, L, flag, None))

    def add(self, scalars, funcs, L, flag=ALL):
        arr = (C.c_void_p * len(funcs))(*[f.h.value for f in funcs])
        check(load().hyg_add(self.h, len(funcs), np.asarray(scalars, dtype=np.float64), arr, L, flag, None))

    def mult(self, a, b, L, flag=ALL):
        check(load().hyg_mult_elementwise(self.h, a.h, b.h, L, flag, None))

    def dot(self, other, L, flag=ALL):
        out = C.c_double()
        check(load().hyg_dot(self.h, other.h, L, flag, C.byref(out), None))
        return out.value

    def convert_from(self, src, L):
        check(load().hyg_convert(self.h, src.h, L, None))

    def restrict(self, L, flag=INNER):
        check(load().hyg_restrict(self.h, L, flag, None))

    def prolongate(self, L, flag=INNER, update=REPLACE):
        check(load().hyg_prolongate(self.h, L, flag, update, None))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hyg_func_destroy(self.h)


class P1ConstantOperator:
    def __init__(self, mesh, minL, maxL):
        self.mesh = mesh
        h = _out()
        check(load().hyg_op_create_p1_const(mesh.h, minL, maxL, C.byref(h)))
        self.h = h

    def cell_stencil(self, L, c):
        out = np.zeros(15)
        check(load().hyg_op_cell_stencil(self.h, L, c, out))
        return out

    def apply(self, src, dst, L, flag=INNER, update=REPLACE):
        check(load().hyg_apply(self.h, src.h, dst.h, L, flag, update, None))

    def smooth_jac(self, dst, rhs, src, relax, L, flag=INNER):
        check(load().hyg_smooth_jac(self.h, dst.h, rhs.h, src.h, relax, L, flag, None))

    def smooth_sor(self, dst, rhs, relax, L, flag=INNER, backwards=False):
        check(load().hyg_smooth_sor(self.h, dst.h, rhs.h, relax, L, flag, 1 if backwards else 0, None))

    def smooth_sor_ordered(self, dst, rhs, relax, L, flag=INNER, backwards=False, stream=None):
        """smooth_sor in the reference's own (lexicographic, primitive by primitive) order."""
        check(load().hyg_smooth_sor_ordered(self.h, dst.h, rhs.h, relax, L, flag, 1 if backwards else 0, stream))

    def inverse_diagonal(self, dst, L):
        check(load().hyg_inverse_diagonal(self.h, dst.h, L, None))

    def residual(self, r, b, x, L, flag=INNER):
        check(load().hyg_residual(self.h, r.h, b.h, x.h, L, flag, None))

    def tune(self, L, dtype=F64, reps=5, stream=None):
        """Time every cell-kernel configuration per stencil mode at level L and keep the fastest."""
        chosen = (C.c_int * 6)()
        check(load().hyg_op_tune(self.h, L, dtype, reps, chosen, stream))
        return list(chosen)

    def config(self, L, dtype=F64, mode=0):
        c, r, p = C.c_int(), C.c_int(), C.c_int()
        check(load().hyg_op_get_config(self.h, L, dtype, mode, C.byref(c), C.byref(r), C.byref(p)))
        return c.value, r.value, p.value

    def set_config(self, L, dtype, mode, cfg):
        check(load().hyg_op_set_config(self.h, L, dtype, mode, cfg))

    def power_iteration(self, x, tmp, iters, L):
        out = C.c_double()
        check(load().hyg_power_iteration(self.h, x.h, tmp.h, iters, L, C.byref(out), None))
        return out.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hyg_op_destroy(self.h)


class P1DivKGradOperator(P1ConstantOperator):
    """-div(k grad u), k = k0 exp(sum_m a_m cos(2 pi kappa_m . x + phi_m)), modes: M x 5 array; or, with
    poly = coefficients of the monomials of degree <= poly_degree (MonomialBasis3D order), the polynomial k."""

    def __init__(self, mesh, minL, maxL, k0=1.0, modes=None, poly=None, poly_degree=None):
        self.mesh = mesh
        h = _out()
        if poly is not None:
            c = np.ascontiguousarray(poly, dtype=np.float64)
            check(load().hyg_op_create_p1_divkgrad_poly(mesh.h, minL, maxL, int(poly_degree), c, C.byref(h)))
        else:
            md = np.zeros((1, 5)) if modes is None or len(modes) == 0 else np.ascontiguousarray(modes, dtype=np.float64)
            nm = 0 if modes is None else len(modes)
            check(load().hyg_op_create_p1_divkgrad(mesh.h, minL, maxL, float(k0), md, nm, C.byref(h)))
        self.h = h

    def store_local_matrices(self):
        """Keep the coefficient integrals of every micro-element (computeAndStoreLocalElementMatrices)."""
        check(load().hyg_op_store_local_matrices(self.h))

    def interpolate_stencils(self, degree, interpolation_level):
        """P1SurrogateOperator::interpolateStencils: polynomial cell-interior stencils from then on."""
        check(load().hyg_op_interpolate_stencils(self.h, degree, interpolation_level))


class P1BlendingLaplaceOperator(P1ConstantOperator):
    """P1ElementwiseBlendingLaplaceOperator on a mesh with the shell map (Mesh.set_shell_map)."""

    def __init__(self, mesh, minL, maxL):
        self.mesh = mesh
        h = _out()
        check(load().hyg_op_create_p1_blending_laplace(mesh.h, minL, maxL, C.byref(h)))
        self.h = h


class P2ConstantOperator(P1ConstantOperator):
    """P2 -Laplace on P2 levels minL..maxL (functions created with space="p2")."""

    def __init__(self, mesh, minL, maxL):
        self.mesh = mesh
        h = _out()
        check(load().hyg_op_create_p2_const(mesh.h, minL, maxL, C.byref(h)))
        self.h = h


class P2DivKGradOperator(P1ConstantOperator):
    """P2 -div(k grad) on P2 levels minL..maxL; kind 0: k = k0 exp(alpha (1 - |x|)) (1 + beta xi_e),
    kind 1: k = 1 + x/2 + y^2/4."""

    def __init__(self, mesh, minL, maxL, kind=0, k0=1.0, alpha=0.0, beta=0.0, seed=0):
        self.mesh = mesh
        h = _out()
        check(load().hyg_op_create_p2_divkgrad(mesh.h, minL, maxL, kind, k0, alpha, beta, seed, C.byref(h)))
        self.h = h

    def store_local_matrices(self):
        """Assemble and keep the per-point stencils (computeAndStoreLocalElementMatrices)."""
        check(load().hyg_op_store_local_matrices(self.h))


def apply_vector(op, srcs, dsts, L, flag=INNER, update=REPLACE, stream=None):
    """The scalar operator applied to each component (VectorLaplaceOperator::apply), up to three at once."""
    n = len(srcs)
    sa = (C.c_void_p * n)(*[f.h.value for f in srcs])
    da = (C.c_void_p * n)(*[f.h.value for f in dsts])
    check(load().hyg_apply_vector(op.h, n, sa, da, L, flag, update, stream))


def cheb_coefficients(order, lower, upper):
    out = np.zeros(5)
    check(load().hyg_cheb_coefficients(order, lower, upper, out))
    return out[:order]


class GMG:
    def __init__(self, op, minL, maxL, dtype=F64, pre=2, post=2, inc=0, smoother="jacobi", relax=2.0 / 3.0,
                 cheb_order=2, cheb_bounds=(0.0, 0.0), cg_max_iter=10000, cg_tol=1e-16):
        self.op, self.maxL = op, maxL
        ip = np.array([minL, maxL, pre, post, inc, {"jacobi": 0, "cheb": 1, "sor": 2, "ssor": 3, "sor_ordered": 4, "ssor_ordered": 5}[smoother], cheb_order, cg_max_iter],
                      dtype=np.int32)
        dp = np.array([relax, cheb_bounds[0], cheb_bounds[1], cg_tol])
        h = _out()
        check(load().hyg_gmg_create(op.h, dtype, ip, dp, C.byref(h)))
        self.h = h

    def vcycle(self, x, b, L):
        check(load().hyg_gmg_vcycle(self.h, x.h, b.h, L, None))

    def vcycle_graph(self, x, b, L, cg_iterations, stream):
        """One V-cycle replayed from a captured hipGraph (captured on the first call)."""
        check(load().hyg_gmg_vcycle_graph(self.h, x.h, b.h, L, cg_iterations, stream))

    def solve(self, x, b, L, cycles):
        res = np.zeros(cycles + 1)
        check(load().hyg_gmg_solve(self.h, x.h, b.h, L, cycles, res, None))
        return res

    def solve_tol(self, x, b, L, max_cycles, rel_tol, stream=None):
        res = np.zeros(max_cycles + 1)
        n = C.c_int()
        check(load().hyg_gmg_solve_tol(self.h, x.h, b.h, L, max_cycles, rel_tol, res, C.byref(n), stream))
        return res[: n.value + 1]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hyg_gmg_destroy(self.h)


def ir_solve(op64, gmg32, x, b, L, inner=1, max_outer=30, rel_tol=1e-10, stream=None):
    res = np.zeros(max_outer + 1)
    steps = C.c_int()
    check(load().hyg_ir_solve(op64.h, gmg32.h, x.h, b.h, L, inner, max_outer, rel_tol, res, C.byref(steps), stream))
    return res[: steps.value + 1]


def _hipdev(nbytes):
    """A device buffer (hipMalloc) for calls that take device pointers; freed by _hipfree."""
    hip = C.CDLL("libamdhip64.so")
    p = C.c_void_p()
    if hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) != 0:
        raise HygError(ERR_HIP, "hipMalloc failed")
    return hip, p


def halo_layout(mesh, L):
    """Element offsets of the reference's per-primitive memory with ghost layers (vertices, edges, faces)."""
    nv, ne, nf = (int(v) for v in mesh.info()[:3])
    off = np.zeros(nv + ne + nf + 1, dtype=np.int64)
    check(load().hyg_halo_layout(mesh.h, L, off))
    return off


def halo_pack(f, L):
    """The reference-layout memory of level L of a P1 function, copied to the host (float64)."""
    n = int(halo_layout(f.mesh, L)[-1])
    es = 8 if f.dtype == F64 else 4
    hip, p = _hipdev(n * es)
    try:
        check(load().hyg_halo_pack(f.h, L, p, None))
        out = np.zeros(n, dtype=np.float64 if es == 8 else np.float32)
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(n * es), 2) != 0:
            raise HygError(ERR_HIP, "hipMemcpy failed")
        return out.astype(np.float64)
    finally:
        hip.hipFree(p)


def halo_unpack(f, L, mem):
    """Own values of every primitive in `mem` (reference layout) into all replicas of the function."""
    es = 8 if f.dtype == F64 else 4
    a = np.ascontiguousarray(mem, dtype=np.float64 if es == 8 else np.float32)
    hip, p = _hipdev(a.nbytes)
    try:
        if hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) != 0:
            raise HygError(ERR_HIP, "hipMemcpy failed")
        check(load().hyg_halo_unpack(f.h, L, p, None))
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(p)


def halo_sum_replicas(f, L, flag=ALL):
    check(load().hyg_halo_sum_replicas(f.h, L, flag, None))


class Batch:
    """Independent problems solved together: every grid operation of their V-cycles is one launch (hyg_batch_*).
    solvers[i] solves problem i on x[i], b[i] at its max level."""

    def __init__(self, solvers, xs, bs):
        n = len(solvers)
        arr = lambda objs: (C.c_void_p * n)(*[o.h.value if isinstance(o.h, C.c_void_p) else o.h for o in objs])
        self._keep = (solvers, xs, bs)
        self.n = n
        h = _out()
        check(load().hyg_batch_create(n, arr(solvers), arr(xs), arr(bs), C.byref(h)))
        self.h = h
        self._arr = arr

    def set_config(self, cfg):
        check(load().hyg_batch_set_config(self.h, cfg))

    def vcycle(self, use_graph=False, stream=None):
        check(load().hyg_batch_vcycle(self.h, 1 if use_graph else 0, stream))

    def apply(self, srcs, dsts, flag=INNER, stream=None):
        check(load().hyg_batch_apply(self.h, self._arr(srcs), self._arr(dsts), flag, stream))

    def residual_norms(self, stream=None):
        out = np.zeros(self.n)
        check(load().hyg_batch_residual_norms(self.h, out, stream))
        return out

    def ir_solve(self, x64, b64, inner=1, max_outer=30, rel_tol=1e-10, use_graph=False, stream=None):
        """Defect correction of every problem; returns the residual histories (n x (steps + 1))."""
        res = np.zeros(self.n * (max_outer + 1))
        steps = C.c_int()
        check(load().hyg_batch_ir_solve(self.h, self._arr(x64), self._arr(b64), inner, max_outer, rel_tol,
                                        1 if use_graph else 0, res, C.byref(steps), stream))
        return res.reshape(self.n, max_outer + 1)[:, : steps.value + 1]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hyg_batch_destroy(self.h)


def fourier_field(seed, modes=32, kmax=8, std=1.0):
    """Mode table of a seeded log-normal coefficient k = k0 exp(g), g = sum_m a cos(2 pi kappa_m . x + phi_m):
    rows (kappa_x, kappa_y, kappa_z, phi, a) with integer kappa, 1 <= max-norm(kappa) <= kmax,
    phi ~ U[0, 2 pi) and a = std sqrt(2 / modes), so that g has standard deviation std."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < modes:
        kap = rng.integers(-kmax, kmax + 1, 3)
        if np.abs(kap).max() == 0:
            continue
        out.append([kap[0], kap[1], kap[2], rng.uniform(0, 2 * np.pi), std * np.sqrt(2.0 / modes)])
    return np.array(out, dtype=np.float64)
