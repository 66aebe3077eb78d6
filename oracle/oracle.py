"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker. The product path (eden_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        for sfx in ("f32", "f32libm", "f32ofast", "f64"):
            run = getattr(L, f"oracle_{sfx}_run")
            run.restype = C.c_int64
            run.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_int64, C.POINTER(C.c_int64), C.c_void_p]
            h = getattr(L, f"oracle_{sfx}_hash64shift")
            h.restype = C.c_uint64
            h.argtypes = [C.c_uint64]
            r = getattr(L, f"oracle_{sfx}_randof")
            r.restype = C.c_float
            r.argtypes = [C.c_float, C.c_int32, C.c_int64, C.c_int64, C.c_int32]
            rt = getattr(L, f"oracle_{sfx}_rate")
            rt.restype = C.c_double
            rt.argtypes = [C.c_void_p, C.c_double]
            sl = getattr(L, f"oracle_{sfx}_set_synapse_log")
            sl.restype = None
            sl.argtypes = [C.c_void_p, C.c_int64]
            sc = getattr(L, f"oracle_{sfx}_synapse_log_count")
            sc.restype = C.c_int64
            sc.argtypes = []
            st = getattr(L, f"oracle_{sfx}_set_trace64")
            st.restype = None
            st.argtypes = [C.c_void_p]
            pl = getattr(L, f"oracle_{sfx}_set_poisson_log")
            pl.restype = None
            pl.argtypes = [C.c_void_p, C.c_int64]
            pc = getattr(L, f"oracle_{sfx}_poisson_log_count")
            pc.restype = C.c_int64
            pc.argtypes = []
            for nm in ("set_lems", "set_lems_cells", "set_lems_rates"):
                sl = getattr(L, f"oracle_{sfx}_{nm}")
                sl.restype = None
                sl.argtypes = [C.c_void_p]
            o = getattr(L, f"oracle_{sfx}_dfs_order")
            o.restype = None
            o.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def run_with_synapse_log(model, n_rows=-1, precision="f32", max_syn_events=10_000_000, **kw):
    """run() plus the synaptic event log as sorted (step, synapse) pairs."""
    L = lib()
    buf = np.zeros(max_syn_events, np.int64)
    getattr(L, f"oracle_{precision}_set_synapse_log")(buf.ctypes.data, max_syn_events)
    try:
        out = run(model, n_rows=n_rows, precision=precision, **kw)
        n = int(getattr(L, f"oracle_{precision}_synapse_log_count")())
    finally:
        getattr(L, f"oracle_{precision}_set_synapse_log")(None, 0)
    if n > max_syn_events:
        raise RuntimeError("synapse log overflow")
    ev = np.sort(buf[:n])
    return out, (ev >> 32, ev & 0xFFFFFFFF)


def run_with_poisson_log(model, n_rows=-1, precision="f32", max_events=10_000_000, library=None, **kw):
    """run() plus the Poisson inputs' firings as sorted (step, input index) pairs. `library`:
    another build of the oracle's source (the tests' planted-mistake variants)."""
    L = library or lib()
    buf = np.zeros(max_events, np.int64)
    getattr(L, f"oracle_{precision}_set_poisson_log")(buf.ctypes.data, max_events)
    try:
        out = run(model, n_rows=n_rows, precision=precision, library=L, **kw)
        n = int(getattr(L, f"oracle_{precision}_poisson_log_count")())
    finally:
        getattr(L, f"oracle_{precision}_set_poisson_log")(None, 0)
    if n > max_events:
        raise RuntimeError("poisson log overflow")
    ev = np.sort(buf[:n])
    return out, (ev >> 32, ev & 0xFFFFFFFF)


def load_variant(path):
    """A separately built copy of the oracle (same source, other macros), with the argument
    types of the f32 entry points set."""
    L = C.CDLL(path)
    run_ = L.oracle_f32_run
    run_.restype = C.c_int64
    run_.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                     C.c_int64, C.POINTER(C.c_int64), C.c_void_p]
    L.oracle_f32_set_poisson_log.restype = None
    L.oracle_f32_set_poisson_log.argtypes = [C.c_void_p, C.c_int64]
    L.oracle_f32_poisson_log_count.restype = C.c_int64
    L.oracle_f32_poisson_log_count.argtypes = []
    return L


# random(x) of a point neuron's step: the oracle's own randof (set after loading), stamped with
# the cell's seed, the step, the call's number and the reference's NOT_AN_INSTANCE marker
# (~0xFee1600dLL, eden/Eden.cpp:3255-3256) as the instance
_RANDOF_PRELUDE = (
    "float (*lems_randof_impl)(float, int, long long, long long, int) = 0;\n"
    "static float lems_randof(float x, int seed, long long step, int n) {\n"
    "    return lems_randof_impl(x, seed, (long long)~0xFee1600dLL, step, n);\n}\n")

_LEMS_PRELUDE = {
    # the oracle's own exp/log: rounded once from a double evaluation (as the f32 flavour)
    "exact": "#include <math.h>\nstatic inline float EXP(float x) { return (float)exp((double)x); }\n"
             "static inline float LOG(float x) { return (float)log((double)x); }\n"
             "static inline float POW(float x, float y) { return (float)pow((double)x, (double)y); }\n"
             + "".join(f"static inline float {f.upper()}(float x) {{ return (float){f}((double)x); }}\n"
                       for f in ("sin", "cos", "tan", "sinh", "cosh", "tanh"))
             + "static float stepf(float x) { if (x < 0) return 0; else return 1; }\n" + _RANDOF_PRELUDE,
    # glibc expf/logf/powf (the f32libm and f32ofast flavours)
    "libm": "#include <math.h>\n#define EXP expf\n#define LOG logf\n#define POW powf\n"
            "#define SIN sinf\n#define COS cosf\n#define TAN tanf\n#define SINH sinhf\n#define COSH coshf\n"
            "#define TANH tanhf\n"
            "static float stepf(float x) { if (x < 0) return 0; else return 1; }\n" + _RANDOF_PRELUDE,
}
_LEMS_FLAGS = {
    "exact": ["-O2", "-ffp-contract=off"],
    "libm": ["-O2", "-ffp-contract=off"],
    "ofast": ["-Ofast", "-march=x86-64-v3", "-mtune=generic"],
}
_lems_libs = {}


def lems_library(source_c, flavour="exact"):
    """The C dialect of a model's generic synapse source compiled with gcc into a shared
    library (cached per source and flavour); returns the ctypes library."""
    import hashlib
    import subprocess
    import tempfile
    text = _LEMS_PRELUDE["exact" if flavour == "exact" else "libm"] + source_c
    key = hashlib.sha1((flavour + "\n" + text).encode()).hexdigest()
    if key in _lems_libs:
        return _lems_libs[key]
    d = os.path.join(tempfile.gettempdir(), "eden_oracle_lems")
    os.makedirs(d, exist_ok=True)
    src, so = os.path.join(d, key + ".c"), os.path.join(d, key + ".so")
    if not os.path.exists(so):
        with open(src, "w") as f:
            f.write(text)
        tmp = so + f".{os.getpid()}.tmp"
        subprocess.run(["gcc", "-std=c11", "-fPIC", "-shared", *_LEMS_FLAGS[flavour], src, "-o", tmp, "-lm"], check=True)
        os.replace(tmp, so)
    lib_ = C.CDLL(so)
    C.c_void_p.in_dll(lib_, "lems_randof_impl").value = C.cast(lib().oracle_f32_randof, C.c_void_p).value
    _lems_libs[key] = lib_
    return lib_


def run(model, n_rows=-1, precision="f32", substeps=1, threads=0, max_events=10_000_000, want_trace=True,
        library=None):
    """Run the restatement on a Model. Returns (trace [rows x n_record] float32,
    event columns, event rows, final voltages). With substeps > 1, dt is divided and the
    trace keeps every substeps-th row; event rows are then in fine steps."""
    L = library or lib()
    d = model.desc()
    if getattr(model, "lems", None) is not None:
        flav = {"f32": "exact", "f32libm": "libm", "f32ofast": "ofast"}.get(precision)
        if flav is None:
            raise RuntimeError("generic LEMS synapses: the oracle runs them in its fp32 flavours only")
        fn = C.cast(lems_library(model.lems.source_c, flav).lems_syn, C.c_void_p)
        getattr(L, f"oracle_{precision}_set_lems")(fn)
    if getattr(model, "lems_cells", None) is not None:
        flav = {"f32": "exact", "f32libm": "libm", "f32ofast": "ofast"}.get(precision)
        if flav is None:
            raise RuntimeError("LEMS point neurons: the oracle runs them in its fp32 flavours only")
        fn = C.cast(lems_library(model.lems_cells.source_c, flav).lems_cell, C.c_void_p)
        getattr(L, f"oracle_{precision}_set_lems_cells")(fn)
    if getattr(model, "lems_rates", None) is not None:
        flav = {"f32": "exact", "f32libm": "libm", "f32ofast": "ofast"}.get(precision)
        if flav is None:
            raise RuntimeError("user rate ComponentTypes: the oracle runs them in its fp32 flavours only")
        fn = C.cast(lems_library(model.lems_rates.source_c, flav).lems_rate, C.c_void_p)
        getattr(L, f"oracle_{precision}_set_lems_rates")(fn)
    total = model.total_rows() if n_rows < 0 else n_rows
    nrec = len(model.record_comp)
    trace = np.zeros((max(total, 1), max(nrec, 1)), np.float32) if want_trace else None
    ev_col = np.zeros(max_events, np.int64)
    ev_row = np.zeros(max_events, np.int64)
    n_ev = C.c_int64()
    final_v = np.zeros(model.n_comps, np.float32)
    fn = getattr(L, f"oracle_{precision}_run")
    rows = fn(C.byref(d), n_rows, substeps, threads,
              None if trace is None else trace.ctypes.data, ev_col.ctypes.data, ev_row.ctypes.data, max_events,
              C.byref(n_ev), final_v.ctypes.data)
    if rows < 0:
        raise RuntimeError("oracle run failed")
    ne = min(n_ev.value, max_events)
    tr = None if trace is None else trace[:rows, :nrec]
    return tr, ev_col[:ne].copy(), ev_row[:ne].copy(), final_v


def run_f64_trace(model, threads=0):
    """FP64 flavour's trace in double precision ([rows x n_record] float64)."""
    L = lib()
    total = model.total_rows()
    nrec = len(model.record_comp)
    buf = np.zeros((max(total, 1), max(nrec, 1)), np.float64)
    L.oracle_f64_set_trace64(buf.ctypes.data)
    try:
        run(model, precision="f64", threads=threads, want_trace=True)
    finally:
        L.oracle_f64_set_trace64(None)
    return buf[:total, :nrec]


def hash64shift(x):
    return int(lib().oracle_f32_hash64shift(x))


def randof(x, work_item, instance, step, invocation):
    return float(lib().oracle_f32_randof(x, work_item, instance, step, invocation))


def rate(r, v, precision="f32"):
    return float(getattr(lib(), f"oracle_{precision}_rate")(C.byref(r), v))


def dfs_order(parent):
    parent = np.ascontiguousarray(parent, np.int32)
    out = np.zeros(len(parent), np.int32)
    lib().oracle_f32_dfs_order(len(parent), parent.ctypes.data, out.ctypes.data)
    return out
