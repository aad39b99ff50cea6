"""Python binding of the serving loop (librvs_host.so, include/rvs_host.h):
resident source sets, target-view requests admitted by memory budget and
batched on arrival (continuous batching), open- or closed-loop runs with
per-request latency.  The loop itself is native (rvs::Server)."""
import ctypes as C

import numpy as np

import rvs_amd as R


class RvsServeOpts(C.Structure):
    _fields_ = [("memory_budget", C.c_size_t), ("max_batch", C.c_int), ("slots", C.c_int), ("entry_cap", C.c_uint32),
                ("copy_to_host", C.c_int), ("keep_outputs", C.c_int), ("host_staging_bytes", C.c_size_t),
                ("compute_streams", C.c_int), ("split_finish", C.c_int)]


class RvsRequest(C.Structure):
    _fields_ = [("scene", C.c_int), ("target", R.RvsTarget)]


class RvsRequestStats(C.Structure):
    _fields_ = [("arrival_ms", C.c_double), ("launch_ms", C.c_double), ("done_ms", C.c_double), ("batch", C.c_int),
                ("batch_views", C.c_int), ("rc", C.c_int)]


def _lib():
    L = R.host()
    if not getattr(L, "_rvs_serve_sig", False):
        PS = C.POINTER(R.RvsSource)
        L.rvs_server_create.argtypes = [C.POINTER(PS), C.POINTER(C.c_int), C.c_int, C.POINTER(RvsServeOpts)]
        L.rvs_server_create.restype = C.c_void_p
        L.rvs_server_destroy.argtypes = [C.c_void_p]
        L.rvs_server_destroy.restype = None
        L.rvs_server_run_open.argtypes = [C.c_void_p, C.POINTER(RvsRequest), C.c_int, C.c_double, C.c_uint64,
                                          C.POINTER(RvsRequestStats)]
        L.rvs_server_run_open.restype = C.c_int
        L.rvs_server_run_closed.argtypes = [C.c_void_p, C.POINTER(RvsRequest), C.c_int, C.c_int,
                                            C.POINTER(RvsRequestStats)]
        L.rvs_server_run_closed.restype = C.c_int
        L.rvs_server_output.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        L.rvs_server_output.restype = C.c_longlong
        L._rvs_serve_sig = True
    return L


def _stats_dict(st):
    d = {k: np.array([getattr(s, k) for s in st]) for k, _ in RvsRequestStats._fields_}
    d["latency_ms"] = d["done_ms"] - d["arrival_ms"]
    return d


class Server:
    """scenes: list of lists of rvs_amd.DeviceSource (kept alive by the server)."""

    def __init__(self, scenes, memory_budget, max_batch=64, slots=2, entry_cap=0, copy_to_host=True,
                 keep_outputs=False, host_staging_bytes=0, compute_streams=1,
                 split_finish=False):
        self._scenes = scenes
        self._arrs = [(R.RvsSource * len(sc))(*[s.struct() for s in sc]) for sc in scenes]
        n = len(scenes)
        ptrs = (C.POINTER(R.RvsSource) * n)(*[C.cast(a, C.POINTER(R.RvsSource)) for a in self._arrs])
        counts = (C.c_int * n)(*[len(sc) for sc in scenes])
        self.opts = RvsServeOpts(int(memory_budget), int(max_batch), int(slots), int(entry_cap), int(copy_to_host),
                                 int(keep_outputs), int(host_staging_bytes), int(compute_streams),
                                 int(split_finish))
        self.h = _lib().rvs_server_create(ptrs, counts, n, C.byref(self.opts))
        if not self.h:
            raise RuntimeError(R.host().rvs_host_last_error().decode())

    def close(self):
        if self.h:
            _lib().rvs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _requests(reqs):
        arr = (RvsRequest * len(reqs))()
        for k, (scene, tgt) in enumerate(reqs):
            arr[k].scene = int(scene)
            arr[k].target = tgt
        return arr

    def _run(self, fn, reqs, *args):
        arr = self._requests(reqs)
        st = (RvsRequestStats * len(reqs))()
        rc = fn(self.h, arr, len(reqs), *args, st)
        if rc != 0:
            err = R.RvsError(rc, R.host().rvs_host_last_error().decode())
            err.stats = _stats_dict(st)   # per-request outcomes of the failed run
            raise err
        return _stats_dict(st)

    def run_open(self, reqs, rate, seed=4):
        """reqs: list of (scene index, RvsTarget); Poisson arrivals at `rate`/s."""
        return self._run(_lib().rvs_server_run_open, reqs, float(rate), int(seed))

    def run_closed(self, reqs, concurrency):
        return self._run(_lib().rvs_server_run_closed, reqs, int(concurrency))

    def output(self, k):
        """Host YUV planes (Y, U, V concatenated, bytes) of request k of the last run."""
        n = _lib().rvs_server_output(self.h, int(k), None, 0)
        if n < 0:
            raise RuntimeError("no output kept for request %d" % k)
        buf = (C.c_uint8 * n)()
        _lib().rvs_server_output(self.h, int(k), buf, n)
        return bytes(buf)


def summarize(stats, wall_s=None):
    lat = stats["latency_ms"]
    t_end = float(stats["done_ms"].max())
    t_begin = float(stats["arrival_ms"].min())
    span = (t_end - t_begin) * 1e-3 if wall_s is None else wall_s
    return {"requests": int(len(lat)), "views_s": round(len(lat) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3),
            "mean_batch": round(float(np.mean(stats["batch_views"])), 2),
            "batches": int(len(np.unique(stats["batch"])))}
