"""The allocation-free apply (ps_mob_mult_noalloc) against the context's apply (ps_mob_mult) on the
same positions and forces: FP64 to 1e-12 relative, on two calls in a row (the second on a smaller
system, so the shared scratch array holds the first call's contents) and a third call that repeats
the first bit for bit.  The test's device buffers and stream come from the HIP runtime the library
links; the call goes through the C ABI."""
import numpy as np
import pytest

from polystokes_brownian_amd import configs
from polystokes_brownian_amd.polystokes import Context, PsError, default_params, mob_mult_noalloc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(w):
    pd = configs.params_dict(w)
    x0 = configs.positions(w)
    g = Context(pd)
    g.initial_configuration(x0)
    g.check_dist()
    g.mob()
    f = np.random.default_rng(w["Np"]).normal(size=g.n)
    u_ctx = g.mob_mult(f)
    g.close()
    return pd, x0, f, u_ctx


class _Hip:
    """The HIP runtime the library itself links (device buffers and a stream for the test)."""

    def __init__(self):
        import ctypes as C
        import os
        from polystokes_brownian_amd.polystokes import lib
        lib()   # the library is loaded first, so the soname below resolves to the runtime it uses
        self.C = C
        try:
            self.L = C.CDLL("libamdhip64.so.7")
        except OSError:
            self.L = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.C.c_void_p()
        self.ok(self.L.hipMalloc(self.C.byref(p), a.nbytes))
        self.ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))   # host to device
        return p

    def get(self, p, n):
        out = np.empty(n)
        self.ok(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))   # device to host
        return out


def _noalloc(pd, x0, f):
    h = _Hip()
    x, fd, u = h.put(x0), h.put(f), h.put(np.full(len(f), np.nan))
    st = h.C.c_void_p()
    h.ok(h.L.hipStreamCreate(h.C.byref(st)))
    mob_mult_noalloc(pd, x.value, fd.value, u.value, st.value)
    h.ok(h.L.hipStreamSynchronize(st))
    out = h.get(u, len(f))
    h.ok(h.L.hipStreamDestroy(st))
    for p in (x, fd, u):
        h.ok(h.L.hipFree(p))
    return out


def test_noalloc_apply_matches_the_context_apply():
    big = _case(configs.small(3, 35, 20, 14.0, seed=12))            # 703 beads, box, three tiles
    small = _case(configs.small(2, 3, 8, 6.0, seed=7, chain_HI=1))   # 26 beads, chain_HI
    u1 = _noalloc(*big[:3])
    print("noalloc vs context, first call", rel(u1, big[3]), flush=True)
    assert rel(u1, big[3]) < 1e-12
    u2 = _noalloc(*small[:3])
    print("noalloc vs context, second (smaller) call", rel(u2, small[3]), flush=True)
    assert rel(u2, small[3]) < 1e-12
    u3 = _noalloc(*big[:3])
    assert np.array_equal(u3, u1)


def test_noalloc_rejects_a_system_beyond_its_scratch_array():
    p = default_params(Nm=2000000, Nc=10, Np=2000010, Npoly=100000, Nmono_per_chain=20, beta=0.2, dt=1e-3)
    with pytest.raises(PsError):
        mob_mult_noalloc(p, 1, 1, 1, 0)   # rejected on the host before any device work
