"""C ABI of the HIP engine: the library loads and exports every declared symbol.
CPU only (no GPU calls)."""
import ctypes as ct
import os
import re

from conftest import REPO

HEADER = os.path.join(REPO, "include", "cauchy_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for")))


def test_header_declares_boundary():
    names = declared_functions()
    for n in ("pycauchy_initialize_lti", "pycauchy_initialize_ltv", "pycauchy_single_step_ltiv",
              "pycauchy_single_step_reset", "pycauchy_single_step_shutdown", "mce_estimator_create",
              "mce_estimator_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cauchyfriendly_amd import _native
    lib = ct.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(REPO, "cauchyfriendly_amd")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(root, f)).read()
                for pat in ("import oracle", "from oracle", "liboracle", "libcauchy_ref", "oracle/"):
                    assert pat not in src, f"{f} refers to the oracle ({pat})"


def test_workloads_are_reproducible():
    from cauchyfriendly_amd import workloads as W
    a, b = W.config0(), W.config0()
    assert (a.zs == b.zs).all() and (a.Phis[0] == b.Phis[0]).all()


def test_collect_outputs_copies_in_order_and_frees():
    """mce_collect_outputs (host code): the batched calls' malloc'd outputs, doubles then the
    last (int) output of each estimator, land in order in two buffers; each pointer is freed
    and cleared."""
    import numpy as np
    from cauchyfriendly_amd import _native
    L = _native.lib()
    libc = ct.CDLL("libc.so.6")
    libc.malloc.restype = ct.c_void_p
    libc.malloc.argtypes = [ct.c_size_t]
    sizes = [[2, 0, 3, 1], [1, 4, 0, 2], [0, 0, 0, 0], [5, 1, 1, 3]]
    count, nout = len(sizes), len(sizes[0])
    outs = (ct.c_void_p * (count * nout))()
    sz = (ct.c_int * (count * nout))()
    rng = np.random.default_rng(5)
    want_d, want_i = [], []
    for i in range(count):
        for j in range(nout):
            n = sizes[i][j]
            k = i * nout + j
            sz[k] = n
            if j == nout - 1:
                v = rng.integers(-1000, 1000, n).astype(np.int32)
                want_i.append(v)
            else:
                v = rng.standard_normal(n)
                want_d.append(v)
            outs[k] = libc.malloc(max(v.nbytes, 1))
            ct.memmove(outs[k], v.ctypes.data, v.nbytes)
    want_d, want_i = np.concatenate(want_d), np.concatenate(want_i)
    dbuf = np.full(want_d.size, np.nan)
    ibuf = np.zeros(want_i.size, dtype=np.int32)
    nd = L.mce_collect_outputs(outs, sz, count, nout, dbuf.ctypes.data_as(ct.POINTER(ct.c_double)),
                               ibuf.ctypes.data_as(ct.POINTER(ct.c_int)))
    assert nd == want_d.size
    assert np.array_equal(dbuf, want_d) and np.array_equal(ibuf, want_i)
    assert all(outs[k] is None for k in range(count * nout))


def test_batch_inputs_are_packed_with_pointers_and_sizes():
    import numpy as np
    from cauchyfriendly_amd.pycauchy import _packed
    arrays = [np.array([1.0, 2.0]), np.zeros(0), [3.0, 4.0, 5.0], np.array([[6.0]])]
    cat, ptrs, n = _packed(arrays)
    assert list(n) == [2, 0, 3, 1]
    assert not ptrs[1]
    for k, a in enumerate(arrays):
        a = np.asarray(a, dtype=np.float64).reshape(-1)
        if a.size:
            assert [ptrs[k][i] for i in range(a.size)] == a.tolist()
    _, ptrs0, n0 = _packed([np.zeros(0), np.zeros(0)])
    assert list(n0) == [0, 0] and not ptrs0[0] and not ptrs0[1]


def _split_top(params):
    out, depth, cur = [], 0, ""
    for ch in params:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _header_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"typedef struct.*?\}\s*\w+;", "", text, flags=re.S)
    funcs = {}
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(((?:[^;])*?)\);", text, flags=re.M | re.S):
        funcs[m.group(1)] = [" ".join(p.split()) for p in _split_top(m.group(2)) if p.strip() not in ("", "void")]
    return funcs


def test_swig_interface_covers_every_wrapped_argument():
    """include/pycauchy_mi355x.i (no SWIG here to build it): every function it wraps has a
    typemap for each array, output and callback argument, by the argument names of the C
    header, and every function it ignores exists."""
    iface = open(os.path.join(REPO, "include", "pycauchy_mi355x.i")).read()
    body = re.sub(r"/\*.*?\*/", "", iface, flags=re.S)
    ignored = set(re.findall(r"%ignore\s+(\w+);", body))
    funcs = _header_functions()
    assert ignored <= set(funcs), f"ignored but not declared: {sorted(ignored - set(funcs))}"
    applies = {}
    for kind, targets in re.findall(r"%apply\s*\(([^)]*)\)\s*\{(.*?)\};", body, flags=re.S):
        for t in re.findall(r"\(([^)]*)\)", targets):
            applies[" ".join(t.split())] = " ".join(kind.split())
    outputs = set()
    for targets in re.findall(r"%apply\s+(?:int|double)\s*\*\s*OUTPUT\s*\{(.*?)\};", body, flags=re.S):
        outputs |= {" ".join(t.split()) for t in targets.split(",")}
    callbacks = set(re.findall(r"%typemap\(in\)\s*void\s*\(\*(\w+)\)", body))
    scalar = re.compile(r"^(int|double|bool|long long|void\*) \w+$")
    wrapped = [f for f in funcs if f not in ignored]
    assert "pycauchy_single_step_ltiv" in wrapped and "pycauchy_initialize_nonlin" in wrapped
    for f in wrapped:
        ps = funcs[f]
        k = 0
        while k < len(ps):
            p = ps[k].replace(" *", "*").replace("* ", "* ")
            pair = (p + ", " + ps[k + 1].replace(" *", "*")) if k + 1 < len(ps) else None
            cb = re.match(r"void \(\*(\w+)\)", p)
            if pair is not None and pair in applies:
                k += 2
            elif p in outputs:
                k += 1
            elif cb:
                assert cb.group(1) in callbacks, f"{f}: callback {cb.group(1)} has no typemap"
                k += 1
            elif p == "char* log_dir" or scalar.match(p):
                k += 1
            else:
                raise AssertionError(f"{f}: argument '{ps[k]}' has no typemap")


def test_search_ordering_refuses_more_than_12_states():
    """pycauchy_set_tr_search_idxs_ordering returns 1, as the reference's does, beyond 12 states
    (host code only)."""
    import numpy as np
    from cauchyfriendly_amd import pycauchy
    assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.arange(13, dtype=np.int32)) == 1


def test_search_ordering_refuses_axes_outside_0_to_11():
    """Each entry of the ordering becomes an axis index into the terms' b vectors, so the
    setter refuses a negative axis or one beyond the 12 supported states and leaves the
    ordering unchanged (host code only); an engine also refuses an axis at or beyond its
    own state dimension before it reduces."""
    import numpy as np
    from cauchyfriendly_amd import pycauchy
    default = np.arange(12, dtype=np.int32)
    try:
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.array([0, -1, 2], dtype=np.int32)) == 1
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.array([12, 0], dtype=np.int32)) == 1
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(np.array([2, 0, 1], dtype=np.int32)) == 0
    finally:
        assert pycauchy.pycauchy_set_tr_search_idxs_ordering(default) == 0
