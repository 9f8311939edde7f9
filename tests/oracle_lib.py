"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
MESHES = os.path.join(ROOT, "meshes")

INNER, DIRICHLET, NEUMANN, FREESLIP, ALL = 1, 2, 4, 8, 15

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(ROOT, "oracle", "oracle.cpp")):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        l = C.CDLL(LIB)
        vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        P = np.ctypeslib.ndpointer
        sig = {
            "or_mesh_load": (vp, [C.c_char_p]),
            "or_mesh_load_tags": (vp, [C.c_char_p]),
            "or_mesh_prims": (None, [vp, P(np.int32), P(np.int32)]),
            "or_halo_pack": (None, [vp, i, P(np.float64)]),
            "or_mesh_shell": (vp, [i, i, d, d]),
            "or_mesh_symmetric_cuboid": (vp, [C.POINTER(C.c_double), C.POINTER(C.c_double), i, i, i]),
            "or_mesh_volume": (d, [vp]),
            "or_mesh_free": (None, [vp]),
            "or_mesh_info": (None, [vp, P(np.int64)]),
            "or_ndofs": (i64, [vp, i]),
            "or_dof_coords": (None, [vp, i, P(np.float64), P(np.uint8)]),
            "or_func_new": (vp, [vp, i, i, i]),
            "or_func_free": (None, [vp]),
            "or_func_set": (None, [vp, i, P(np.float64)]),
            "or_func_get": (None, [vp, i, P(np.float64)]),
            "or_op_new": (vp, [vp, i, i, i]),
            "or_op_free": (None, [vp]),
            "or_op_new_divkgrad": (vp, [vp, i, i, i, i, d, d, P(np.float64), i]),
            "or_op_new_divkgrad_poly": (vp, [vp, i, i, i, P(np.float64), i]),
            "or_smooth_sor_ordered": (i, [vp, vp, vp, d, i, i, i]),
            "or_mesh_set_shell_map": (i, [vp]),
            "or_mesh_shell_map": (i, [vp, i, P(np.float64)]),
            "or_op_new_blending_laplace": (vp, [vp, i, i, i]),
            "or_op_interpolate_stencils": (i, [vp, i, i]),
            "or_op_surrogate_poly": (i, [vp, i, i, P(np.float64)]),
            "or_op_variable_stencil": (i, [vp, i, i, C.c_long, C.c_long, C.c_long, P(np.float64)]),
            "or_op_cell_stencil": (None, [vp, i, i, P(np.float64)]),
            "or_apply": (None, [vp, vp, vp, i, i, i]),
            "or_inv_diag": (None, [vp, i, P(np.float64)]),
            "or_smooth_jac": (None, [vp, vp, vp, vp, d, i, i]),
            "or_mesh_cells": (None, [vp, P(np.int32)]),
            "or_mesh_cuboid": (vp, [C.POINTER(C.c_double), C.POINTER(C.c_double), i, i, i]),
            "or_mesh_cubed_domain": (vp, [P(np.int32), i, i]),
            "or_restrict": (None, [vp, i, i]),
            "or_prolongate": (None, [vp, i, i, i]),
            "or_dot": (d, [vp, vp, i, i]),
            "or_cheb_coefficients": (i, [i, d, d, P(np.float64)]),
            "or_power_iteration": (d, [vp, vp, i, i]),
            "or_gmg": (None, [vp, vp, vp, P(np.int32), P(np.float64), i, P(np.float64)]),
            "or_ir": (None, [vp, vp, vp, vp, P(np.int32), P(np.float64), i, i, P(np.float64)]),
            "or_op_new_p2": (vp, [vp, i, i, i]),
            "or_op_new_p2_divkgrad": (vp, [vp, i, i, i, i, d, d, d, C.c_uint64]),
            "or_p2_divkgrad_element_poly": (None, [P(np.float64), P(np.float64)]),
            "or_p2_quadrature": (None, [P(np.float64), P(np.float64)]),
            "or_p2k_value": (d, [i, d, d, d, C.c_uint64, P(np.float64), C.c_uint64]),
            "or_p2_element": (None, [P(np.float64), P(np.float64)]),
            "or_p2_restrict": (None, [vp, i, i]),
            "or_p2_prolongate": (None, [vp, i, i, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class Mesh:
    """A Gmsh v2.2 file (name or path), the spherical shell ("shell", ntan, nrad, rmin, rmax), the symmetric
    cuboid ("symmetric_cuboid", lower xyz, upper xyz, nx, ny, nz) or the six-tet cuboid ("cuboid", same arguments)."""

    def __init__(self, name, *args, tags=False):
        if name == "shell":
            self.path = None
            self.h = lib().or_mesh_shell(*args)
            if not self.h:
                raise RuntimeError("bad spherical shell parameters")
            return
        if name == "cubed_domain":
            self.path = None
            cubes, cube_type = args
            cb = np.ascontiguousarray(np.asarray(cubes, dtype=np.int32).reshape(-1))
            self.h = lib().or_mesh_cubed_domain(cb, len(cb) // 3, cube_type)
            if not self.h:
                raise RuntimeError("bad cubed domain parameters")
            return
        if name in ("symmetric_cuboid", "cuboid"):
            self.path = None
            lo, hi, nx, ny, nz = args
            fn = lib().or_mesh_symmetric_cuboid if name == "symmetric_cuboid" else lib().or_mesh_cuboid
            self.h = fn((C.c_double * 3)(*lo), (C.c_double * 3)(*hi), nx, ny, nz)
            if not self.h:
                raise RuntimeError("bad symmetric cuboid parameters")
            return
        path = name if os.path.isabs(name) else os.path.join(MESHES, name)
        self.path = path
        self.h = (lib().or_mesh_load_tags if tags else lib().or_mesh_load)(path.encode())
        if not self.h:
            raise RuntimeError("cannot load mesh " + path)

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        lib().or_mesh_info(self.h, out)
        return out[:7]

    def ndofs(self, L):
        return lib().or_ndofs(self.h, L)

    def cells(self):
        """Macro-cell vertex indices, (nC, 4), each cell in its local vertex order."""
        nc = int(self.info()[3])
        out = np.zeros(4 * nc, dtype=np.int32)
        lib().or_mesh_cells(self.h, out)
        return out.reshape(nc, 4)

    def prims(self):
        """Edge (nE, 2) and face (nF, 3) vertex lists in storage order."""
        nv, ne, nf = (int(v) for v in self.info()[:3])
        e = np.zeros(2 * ne, dtype=np.int32)
        f = np.zeros(3 * nf, dtype=np.int32)
        lib().or_mesh_prims(self.h, e, f)
        return e.reshape(ne, 2), f.reshape(nf, 3)

    def volume(self):
        return lib().or_mesh_volume(self.h)

    def set_shell_map(self):
        """IcosahedralShellMap::setMap: blended cells; coords() returns blended positions from then on."""
        assert lib().or_mesh_set_shell_map(self.h) == 0

    def shell_map(self, c):
        out = np.zeros(14)
        assert lib().or_mesh_shell_map(self.h, c, out) == 0
        return out[0:3], out[3:6], out[6:9], out[9:12], out[12], out[13]

    def coords(self, L):
        n = self.ndofs(L)
        xyz = np.zeros(3 * n)
        bnd = np.zeros(n, dtype=np.uint8)
        lib().or_dof_coords(self.h, L, xyz, bnd)
        return xyz.reshape(n, 3), bnd.astype(bool)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_mesh_free(self.h)


class Function:
    def __init__(self, mesh, minL, maxL, prec=8):
        self.mesh = mesh
        self.h = lib().or_func_new(mesh.h, minL, maxL, prec)

    def set(self, L, v):
        lib().or_func_set(self.h, L, np.ascontiguousarray(v, dtype=np.float64))

    def get(self, L):
        out = np.zeros(self.mesh.ndofs(L))
        lib().or_func_get(self.h, L, out)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_func_free(self.h)


class Operator:
    def __init__(self, mesh, minL, maxL, prec=8):
        self.mesh = mesh
        self.h = lib().or_op_new(mesh.h, minL, maxL, prec)

    def apply(self, src, dst, L, flag=INNER, add=False):
        lib().or_apply(self.h, src.h, dst.h, L, flag, int(add))

    def cell_stencil(self, L, c):
        out = np.zeros(15)
        lib().or_op_cell_stencil(self.h, L, c, out)
        return out

    def inv_diag(self, L):
        out = np.zeros(self.mesh.ndofs(L))
        lib().or_inv_diag(self.h, L, out)
        return out

    def smooth_jac(self, dst, rhs, src, relax, L, flag=INNER):
        lib().or_smooth_jac(self.h, dst.h, rhs.h, src.h, relax, L, flag)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_op_free(self.h)


class P2Operator(Operator):
    """P2 constant -Laplace on P2 levels minL..maxL. A P2 function of level L is a Function holding lattice
    level L+1 (vertex DoFs at even lattice points, edge DoFs at the midpoints); the apply / inv_diag /
    smooth_jac level arguments are lattice levels."""

    def __init__(self, mesh, minL, maxL, prec=8):
        self.mesh = mesh
        self.h = lib().or_op_new_p2(mesh.h, minL, maxL, prec)


class P2DivKGradOperator(Operator):
    """P2 elementwise -div(k grad) on P2 levels minL..maxL (11-point rule); kind 0: k = k0 exp(alpha (1 - |x|))
    (1 + beta xi_e) with xi_e per micro-element from SplitMix64(seed, element id); kind 1: 1 + x/2 + y^2/4."""

    def __init__(self, mesh, minL, maxL, kind=0, k0=1.0, alpha=0.0, beta=0.0, seed=0, prec=8):
        self.mesh = mesh
        self.h = lib().or_op_new_p2_divkgrad(mesh.h, minL, maxL, prec, kind, k0, alpha, beta, seed)


def p2_divkgrad_element_poly(x):
    """P2 element matrix of -div(k grad) with k = 1 + x/2 + y^2/4 (11-point rule) on tetrahedron x (4 x 3)."""
    out = np.zeros(100)
    lib().or_p2_divkgrad_element_poly(np.ascontiguousarray(x, dtype=np.float64).ravel(), out)
    return out.reshape(10, 10)


def p2_quadrature():
    """The oracle's 11-point rule: barycentric coordinates (11, 4) and weights (11,)."""
    lam, w = np.zeros(44), np.zeros(11)
    lib().or_p2_quadrature(lam, w)
    return lam.reshape(11, 4), w


def p2_element(x):
    """P2 element matrix of -Laplace on the tetrahedron with vertices x (4 x 3)."""
    out = np.zeros(100)
    lib().or_p2_element(np.ascontiguousarray(x, dtype=np.float64).ravel(), out)
    return out.reshape(10, 10)


def p2_restrict(f, src_lat, flag=INNER):
    lib().or_p2_restrict(f.h, src_lat, flag)


def p2_prolongate(f, src_lat, flag=INNER, add=False):
    lib().or_p2_prolongate(f.h, src_lat, flag, int(add))


class BlendingLaplaceOperator(Operator):
    """P1ElementwiseBlendingLaplaceOperator (p1_diffusion_blending_q3) on a mesh with the shell map."""

    def __init__(self, mesh, minL, maxL, prec=8):
        self.mesh = mesh
        self.h = lib().or_op_new_blending_laplace(mesh.h, minL, maxL, prec)


class DivKGradOperator(Operator):
    """P1 elementwise -div(k grad) operator; kfield = ("fourier", k0, modes[M x 5]), ("tanh", alpha) or
    ("poly", coefficients in the monomial basis). interpolate_stencils(degree, level) turns it into the surrogate."""

    def __init__(self, mesh, minL, maxL, kfield, prec=8):
        self.mesh = mesh
        self.degree = -1
        if kfield[0] == "poly":
            c = np.ascontiguousarray(np.asarray(kfield[1], dtype=np.float64))
            self.h = lib().or_op_new_divkgrad_poly(mesh.h, minL, maxL, prec, c, len(c))
        elif kfield[0] == "fourier":
            modes = np.ascontiguousarray(np.asarray(kfield[2], dtype=np.float64).reshape(-1, 5))
            self.h = lib().or_op_new_divkgrad(mesh.h, minL, maxL, prec, 0, float(kfield[1]), 0.0,
                                               modes if len(modes) else np.zeros((1, 5)), len(modes))
        else:
            self.h = lib().or_op_new_divkgrad(mesh.h, minL, maxL, prec, 1, 1.0, float(kfield[1]), np.zeros((1, 5)), 0)

    def interpolate_stencils(self, degree, level):
        assert lib().or_op_interpolate_stencils(self.h, degree, level) == 0
        self.degree = degree

    def surrogate_poly(self, L, c):
        nc = (self.degree + 1) * (self.degree + 2) * (self.degree + 3) // 6
        out = np.zeros(15 * nc)
        assert lib().or_op_surrogate_poly(self.h, L, c, out) == 0
        return out.reshape(15, nc)

    def variable_stencil(self, L, c, x, y, z):
        out = np.zeros(15)
        assert lib().or_op_variable_stencil(self.h, L, c, x, y, z, out) == 0
        return out


def random_field(seed, modes=32, kmax=8, std=1.0):
    """Fourier modes of k = exp(g), g = sum_m a cos(2 pi kappa_m.x + phi_m), integer kappa with
    1 <= max-norm <= kmax, phi ~ U[0, 2 pi), a = std sqrt(2 / modes) (g has standard deviation std)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < modes:
        kap = rng.integers(-kmax, kmax + 1, 3)
        if np.abs(kap).max() == 0:
            continue
        out.append([kap[0], kap[1], kap[2], rng.uniform(0, 2 * np.pi), std * np.sqrt(2.0 / modes)])
    return np.array(out, dtype=np.float64)


def cheb_coefficients(order, lower, upper):
    out = np.zeros(order)
    assert lib().or_cheb_coefficients(order, lower, upper, out) == 0
    return out


def gmg(op, x, b, minL, maxL, cycles, pre=2, post=2, inc=0, smoother="jacobi", relax=2.0 / 3.0,
        cheb_order=2, cheb_bounds=(0.0, 0.0), cg_max_iter=10000, cg_tol=1e-16):
    kind = {"jacobi": 0, "cheb": 1, "sor_ordered": 4, "ssor_ordered": 5}[smoother]
    ip = np.array([minL, maxL, pre, post, inc, kind, cheb_order, cg_max_iter], dtype=np.int32)
    dp = np.array([relax, cheb_bounds[0], cheb_bounds[1], cg_tol])
    res = np.zeros(cycles + 1)
    lib().or_gmg(op.h, x.h, b.h, ip, dp, cycles, res)
    return res
