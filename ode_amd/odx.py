"""ctypes binding of libodx.so (the HIP step path) and the batched-worlds API.

The library is built in-tree by `python -m ode_amd.build`.  Loading fails
loudly when it is missing; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libodx.so")

BODY = 32
GEOM = 16
JOINT = 16
ROW = 32

_lib = None


class OdxError(RuntimeError):
    pass


_MSG_CB_T = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_char_p)
_messages = []


@_MSG_CB_T
def _on_message(err, msg):
    _messages.append((err, msg.decode(errors="replace")))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OdxError("libodx.so is not built (run `python -m ode_amd.build`): " + LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, fp, dp, i = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int
        L.odxVersion.restype = ctypes.c_char_p
        L.odxBatchCreate.restype = ctypes.c_void_p
        L.odxBatchCreate.argtypes = [i, ip, fp, ip, fp, ip, fp, dp, i]
        L.odxBatchCreateJoints.restype = ctypes.c_void_p
        L.odxBatchCreateJoints.argtypes = [i, ip, fp, ip, fp, ip, fp, fp, dp, i]
        L.odxBatchDestroy.argtypes = [vp]
        L.odxBatchSetSolver.argtypes = [vp, i, i, ctypes.c_float, i, i]
        L.odxBatchSetSolveOptions.argtypes = [vp, i, i]
        L.odxBatchSetSurfaceTable.argtypes = [vp, i, fp, i]
        L.odxBatchSaveSnapshot.argtypes = [vp, ctypes.c_char_p]
        L.odxBatchLoadSnapshot.restype = ctypes.c_void_p
        L.odxBatchLoadSnapshot.argtypes = [ctypes.c_char_p]
        L.odxBatchGetLayout.argtypes = [vp, ip, ip, ip]
        L.odxBatchSetSplits.argtypes = [vp, i, i]
        L.odxBatchSetFused.argtypes = [vp, i]
        L.odxBatchAutotune.argtypes = [vp, i, i, ip, fp]
        L.odxBatchStep.argtypes = [vp, i]
        L.odxBatchSync.argtypes = [vp]
        L.odxBatchTime.argtypes = [vp, i, fp]
        L.odxBatchKernelTimes.argtypes = [vp, i, fp]
        L.odxBatchScheduleKernelTimes.argtypes = [vp, i, fp, ip]
        L.odxBatchGetBodies.argtypes = [vp, fp]
        L.odxBatchSetBodies.argtypes = [vp, fp]
        L.odxBatchGetStats.argtypes = [vp, ip, ip, ip, fp]
        L.odxBatchGetJoints.argtypes = [vp, i, fp, i]
        L.odxBatchGetLambda.argtypes = [vp, i, fp, i]
        L.odxBatchGetRows.argtypes = [vp, i, fp, i]
        L.odxBatchGetOverflow.argtypes = [vp, ip]
        L.odxBatchGetSolveTimes.argtypes = [vp, dp]
        L.odxBatchGetPrepareTimes.argtypes = [vp, dp]
        L.odxBatchRowStarts.argtypes = [vp, ip]
        L.odxBatchSetLambda0.argtypes = [vp, fp, i]
        L.odxLargeCreate.restype = ctypes.c_void_p
        L.odxLargeCreate.argtypes = [i, fp, i, fp, dp, i, i]
        L.odxLargeDestroy.argtypes = [vp]
        for name in ("odxLargeSetIterations", "odxLargeStep", "odxLargeSetSolveBlocks", "odxLargeSetSpinLimit"):
            getattr(L, name).argtypes = [vp, i]
        L.odxLargeSync.argtypes = [vp]
        L.odxLargeSetSurfaceTable.argtypes = [vp, i, fp, i]
        for name in ("odxLargeTime", "odxLargeKernelTimes"):
            getattr(L, name).argtypes = [vp, i, fp]
        for name in ("odxLargeGetBodies", "odxLargeSetBodies", "odxLargeGetCounts"):
            getattr(L, name).argtypes = [vp, fp]
        for name in ("odxLargeGetPairs", "odxLargeGetJoints", "odxLargeGetRows", "odxLargeGetLambda",
                     "odxLargeGetColours"):
            getattr(L, name).argtypes = [vp, fp, i]
        L.odxSetMessageHandler(_on_message)
        _lib = L
    return _lib


def last_messages(clear=True):
    out = list(_messages)
    if clear:
        _messages.clear()
    return out


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(ok, what):
    if not ok:
        raise OdxError("%s failed: %s" % (what, "; ".join(m for _, m in last_messages()) or "unknown error"))


def params_vector(h=0.01, gravity=(0, 0, -9.81), erp=0.2, cfm=1e-5, w=1.3, iters=20, max_vel=np.inf, min_depth=0.0,
                  mode=0x7000, mu=0.5, mu2=0.0, bounce=0.0, bounce_vel=0.0, soft_erp=0.0, soft_cfm=0.0, maxc=4,
                  space=0, skip_connected=0, lambda_floor=1e-3, row_order=0):
    """PARAMS vector of include/odx.h (dWorldCreate defaults for dSINGLE).
    row_order (slot 31) is the large world's sweep order: 0 ReorderPrep's,
    2 each contact's rows together."""
    p = np.zeros(32, np.float64)
    p[0] = h
    p[1:4] = gravity
    p[4], p[5], p[6], p[7] = erp, cfm, w, iters
    p[8], p[9] = max_vel, min_depth
    p[10], p[11], p[12], p[13], p[14], p[15], p[16] = mode, mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm
    p[17], p[18] = maxc, space
    p[24], p[25] = skip_connected, lambda_floor
    p[31] = row_order
    return p


def pack_worlds(worlds):
    """List of scenes.World -> concatenated record arrays + prefix offsets."""
    bs, gs, ls = [0], [0], [0]
    for w in worlds:
        bs.append(bs[-1] + len(w.bodies))
        gs.append(gs[-1] + len(w.geoms))
        ls.append(ls[-1] + len(w.balls))
    bodies = np.ascontiguousarray(np.vstack([w.bodies for w in worlds]), np.float32)
    geoms = np.ascontiguousarray(np.vstack([w.geoms for w in worlds]), np.float32)
    balls = [w.balls for w in worlds if len(w.balls)]
    balls = np.ascontiguousarray(np.vstack(balls), np.float32) if balls else np.zeros((0, JOINT), np.float32)
    return (np.asarray(bs, np.int32), bodies, np.asarray(gs, np.int32), geoms, np.asarray(ls, np.int32), balls)


class Batch:
    """Many independent worlds stepped together on the GPU (odxBatch*)."""

    ORDER_REFERENCE = 0
    ORDER_COLOURED = 1

    def __init__(self, worlds, params, contacts_per_body=8):
        L = lib()
        self.worlds = worlds
        bs, bodies, gs, geoms, ls, balls = pack_worlds(worlds)
        self.body_start, self.geom_start, self.ball_start = bs, gs, ls
        self.nworlds = len(worlds)
        self.nbodies = int(bs[-1])
        self.params = np.ascontiguousarray(params, np.float64)
        # extended records of hinge / slider / fixed joints (World.jext, one per
        # joint record; zeros for ball joints)
        jext = None
        if any(getattr(w, "jext", None) is not None for w in worlds):
            jext = np.zeros((len(balls), 32), np.float32)
            for wi, w in enumerate(worlds):
                if getattr(w, "jext", None) is not None and len(w.balls):
                    jext[ls[wi]:ls[wi + 1]] = np.asarray(w.jext, np.float32).reshape(-1, 32)
        self._keep = (bodies, geoms, balls, jext)
        if jext is None:
            h = L.odxBatchCreate(self.nworlds, _p(bs), _p(bodies), _p(gs), _p(geoms), _p(ls), _p(balls),
                                 _p(self.params), contacts_per_body)
        else:
            h = L.odxBatchCreateJoints(self.nworlds, _p(bs), _p(bodies), _p(gs), _p(geoms), _p(ls), _p(balls),
                                       _p(jext), _p(self.params), contacts_per_body)
        _check(h, "odxBatchCreate")
        self.h = h
        rs = np.zeros(self.nworlds + 1, np.int32)
        L.odxBatchRowStarts(self.h, _p(rs))
        self.row_start = rs

    @classmethod
    def load_snapshot(cls, path):
        """A batch re-created from odxBatchSaveSnapshot's file."""
        L = lib()
        h = L.odxBatchLoadSnapshot(str(path).encode())
        _check(h, "odxBatchLoadSnapshot")
        self = cls.__new__(cls)
        self.h = h
        nw, nb = ctypes.c_int(0), ctypes.c_int(0)
        L.odxBatchGetLayout(h, ctypes.byref(nw), ctypes.byref(nb), None)
        self.nworlds, self.nbodies = nw.value, nb.value
        self.body_start = np.zeros(self.nworlds + 1, np.int32)
        L.odxBatchGetLayout(h, None, None, _p(self.body_start))
        self.worlds = None
        rs = np.zeros(self.nworlds + 1, np.int32)
        L.odxBatchRowStarts(h, _p(rs))
        self.row_start = rs
        return self

    def save_snapshot(self, path):
        _check(lib().odxBatchSaveSnapshot(self.h, str(path).encode()), "odxBatchSaveSnapshot")

    def close(self):
        if getattr(self, "h", None):
            lib().odxBatchDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solver(self, order=0, max_iters=20, tol=0.0, check_every=1, warm=0):
        _check(lib().odxBatchSetSolver(self.h, order, max_iters, float(tol), check_every, warm), "odxBatchSetSolver")

    def set_splits(self, nsplit=1, graph_steps=1):
        """Step nsplit world ranges on streams of their own, graph_steps steps
        per graph launch (results equal those of one range)."""
        _check(lib().odxBatchSetSplits(self.h, nsplit, graph_steps), "odxBatchSetSplits")

    def set_fused(self, steps_per_launch):
        """> 0: that many complete steps per launch of the persistent fused
        kernel (0: staged launches)."""
        _check(lib().odxBatchSetFused(self.h, steps_per_launch), "odxBatchSetFused")

    def autotune(self, steps=10, reps=2):
        """Times the launch schedules (splits 1/2/4/8, staged or fused launches
        of the current graph size) over `steps` steps from the current state
        and keeps the fastest; returns
        ({"splits", "graph_steps", "fused"}, ms per step)."""
        sch = np.zeros(3, np.int32)
        ms = np.zeros(1, np.float32)
        _check(lib().odxBatchAutotune(self.h, steps, reps, _p(sch), _p(ms)), "odxBatchAutotune")
        return {"splits": int(sch[0]), "graph_steps": int(sch[1]), "fused": int(sch[2])}, float(ms[0])

    def set_solve_options(self, restart_after=0, lds_rows=400):
        """restart_after: sweeps after which an unconverged warm start restarts
        cold (0 = never); lds_rows: row budget of the LDS-resident sweeps."""
        _check(lib().odxBatchSetSolveOptions(self.h, restart_after, lds_rows), "odxBatchSetSolveOptions")

    def set_surface_table(self, table, category_filter=False):
        """Near-callback policy on the device: table [nmat, nmat, 16] (rows as
        include/odx.h's odxBatchSetSurfaceTable), geoms' materials in slot 13
        and category / collide bits in slots 14 / 15; None returns to the
        parameters' surface."""
        if table is None:
            _check(lib().odxBatchSetSurfaceTable(self.h, 0, None, 0), "odxBatchSetSurfaceTable")
            return
        t = np.ascontiguousarray(table, np.float32)
        _check(lib().odxBatchSetSurfaceTable(self.h, t.shape[0], _p(t), int(category_filter)),
               "odxBatchSetSurfaceTable")

    def step(self, n=1):
        _check(lib().odxBatchStep(self.h, n), "odxBatchStep")

    def sync(self):
        _check(lib().odxBatchSync(self.h), "odxBatchSync")

    def time(self, n):
        out = np.zeros(1, np.float32)
        _check(lib().odxBatchTime(self.h, n, _p(out)), "odxBatchTime")
        return float(out[0])

    def kernel_times(self, n):
        """Average [prepare, warm-start inference, solve] kernel milliseconds."""
        out = np.zeros(3, np.float32)
        _check(lib().odxBatchKernelTimes(self.h, n, _p(out)), "odxBatchKernelTimes")
        return float(out[0]), float(out[1]), float(out[2])

    def schedule_kernel_times(self, n):
        """Per world range of the current schedule: (w0, w1, [prepare,
        inference, solve] mean milliseconds per launch), timed on the range's
        own stream."""
        ms = np.zeros(3 * 8, np.float32)
        rg = np.zeros(2 * 8, np.int32)
        k = lib().odxBatchScheduleKernelTimes(self.h, n, _p(ms), _p(rg))
        _check(k, "odxBatchScheduleKernelTimes")
        return [(int(rg[2 * r]), int(rg[2 * r + 1]), [float(x) for x in ms[3 * r:3 * r + 3]]) for r in range(k)]

    def bodies(self):
        out = np.zeros((self.nbodies, BODY), np.float32)
        _check(lib().odxBatchGetBodies(self.h, _p(out)), "odxBatchGetBodies")
        return out

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, np.float32)
        _check(lib().odxBatchSetBodies(self.h, _p(b)), "odxBatchSetBodies")

    def world_bodies(self, w, bodies=None):
        b = self.bodies() if bodies is None else bodies
        return b[self.body_start[w]:self.body_start[w + 1]]

    def stats(self):
        n = self.nworlds
        nc, nr, ni = (np.zeros(n, np.int32) for _ in range(3))
        res = np.zeros(n, np.float32)
        _check(lib().odxBatchGetStats(self.h, _p(nc), _p(nr), _p(ni), _p(res)), "odxBatchGetStats")
        return {"contacts": nc, "rows": nr, "iters": ni, "residual": res}

    def joints(self, w, cap=8192):
        out = np.zeros((cap, JOINT), np.float32)
        n = lib().odxBatchGetJoints(self.h, w, _p(out), cap)
        return out[:n]

    def lam(self, w, cap=8192):
        out = np.zeros(cap, np.float32)
        n = lib().odxBatchGetLambda(self.h, w, _p(out), cap)
        return out[:n]

    def rows(self, w, cap=8192):
        out = np.zeros((cap, ROW), np.float32)
        n = lib().odxBatchGetRows(self.h, w, _p(out), cap)
        return out[:n]

    def overflow(self):
        out = np.zeros(4, np.int32)
        _check(lib().odxBatchGetOverflow(self.h, _p(out)), "odxBatchGetOverflow")
        return out

    def solve_times(self):
        """[nworlds, 2] start/end of each world's solve in the last step (us)."""
        out = np.zeros((self.nworlds, 2), np.float64)
        _check(lib().odxBatchGetSolveTimes(self.h, _p(out)), "odxBatchGetSolveTimes")
        return out

    def prepare_times(self):
        """[nworlds, 6] prepare-stage boundaries in the last step (us): start,
        bodies, broadphase, narrowphase, rows, schedule."""
        out = np.zeros((self.nworlds, 6), np.float64)
        _check(lib().odxBatchGetPrepareTimes(self.h, _p(out)), "odxBatchGetPrepareTimes")
        return out

    def set_lambda0(self, lam0):
        a = np.ascontiguousarray(lam0, np.float32)
        _check(lib().odxBatchSetLambda0(self.h, _p(a), len(a)), "odxBatchSetLambda0")


class Large:
    """One large world (hash-space broadphase, colour-parallel QuickStep over
    the whole GPU; odxLarge*)."""

    COUNTERS = ("pairs", "contacts", "rows", "colours", "pair_overflow", "contact_overflow", "uncoloured",
                "barrier_error", "colour_rounds")

    def __init__(self, world, params, pair_cap=0, contact_cap=0):
        L = lib()
        self.bodies0 = np.ascontiguousarray(world.bodies, np.float32)
        self.geoms = np.ascontiguousarray(world.geoms, np.float32)
        self.nbodies, self.ngeoms = len(self.bodies0), len(self.geoms)
        self.params = np.ascontiguousarray(params, np.float64)
        h = L.odxLargeCreate(self.nbodies, _p(self.bodies0), self.ngeoms, _p(self.geoms), _p(self.params), pair_cap,
                             contact_cap)
        _check(h, "odxLargeCreate")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().odxLargeDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_iterations(self, iters):
        _check(lib().odxLargeSetIterations(self.h, iters), "odxLargeSetIterations")

    def set_solve_blocks(self, per_cu):
        """Workgroups per CU of the cooperative solve; returns the grid size."""
        b = lib().odxLargeSetSolveBlocks(self.h, per_cu)
        _check(b, "odxLargeSetSolveBlocks")
        return b

    def set_spin_limit(self, polls=1 << 21):
        """Polls one grid-barrier wait may make before the step fails."""
        _check(lib().odxLargeSetSpinLimit(self.h, polls), "odxLargeSetSpinLimit")

    def set_surface_table(self, table, category_filter=False):
        """The material-pair surface table and category / collide bit test of
        Batch.set_surface_table for the large world; None returns to the
        parameters' surface."""
        if table is None:
            _check(lib().odxLargeSetSurfaceTable(self.h, 0, None, 0), "odxLargeSetSurfaceTable")
            return
        t = np.ascontiguousarray(table, np.float32)
        _check(lib().odxLargeSetSurfaceTable(self.h, t.shape[0], _p(t), int(category_filter)),
               "odxLargeSetSurfaceTable")

    def step(self, n=1):
        _check(lib().odxLargeStep(self.h, n), "odxLargeStep")

    def sync(self):
        _check(lib().odxLargeSync(self.h), "odxLargeSync")

    def time(self, n):
        out = np.zeros(1, np.float32)
        _check(lib().odxLargeTime(self.h, n, _p(out)), "odxLargeTime")
        return float(out[0])

    def kernel_times(self, n):
        """Average [prepare, colouring, solve] milliseconds per step."""
        out = np.zeros(3, np.float32)
        _check(lib().odxLargeKernelTimes(self.h, n, _p(out)), "odxLargeKernelTimes")
        return float(out[0]), float(out[1]), float(out[2])

    def bodies(self):
        out = np.zeros((self.nbodies, BODY), np.float32)
        _check(lib().odxLargeGetBodies(self.h, _p(out)), "odxLargeGetBodies")
        return out

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, np.float32)
        _check(lib().odxLargeSetBodies(self.h, _p(b)), "odxLargeSetBodies")

    def counts(self):
        out = np.zeros(16, np.int32)
        _check(lib().odxLargeGetCounts(self.h, _p(out)), "odxLargeGetCounts")
        return {k: int(out[i]) for i, k in enumerate(self.COUNTERS)}

    def _get(self, fn, shape, dtype, n):
        out = np.zeros((max(n, 1),) + shape, dtype)
        k = getattr(lib(), fn)(self.h, _p(out), n)
        return out[:k]

    def pairs(self):
        return self._get("odxLargeGetPairs", (2,), np.int32, self.counts()["pairs"])

    def joints(self):
        return self._get("odxLargeGetJoints", (JOINT,), np.float32, self.counts()["contacts"])

    def rows(self):
        return self._get("odxLargeGetRows", (ROW,), np.float32, self.counts()["rows"])

    def lam(self):
        return self._get("odxLargeGetLambda", (), np.float32, self.counts()["rows"])

    def colours(self):
        return self._get("odxLargeGetColours", (), np.int32, self.counts()["contacts"])


NPARAM = 19075
FEAT = 16


def _setup_mlp_api(L):
    vp, i, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p
    L.odxMlpCreate.restype = ctypes.c_void_p
    L.odxMlpCreate.argtypes = [ctypes.c_uint]
    L.odxMlpDestroy.argtypes = [vp]
    L.odxMlpSetNormalization.argtypes = [vp, fp]
    L.odxMlpSetGate.argtypes = [vp, i]
    L.odxMlpGetParams.argtypes = [vp, fp]
    L.odxMlpSetParams.argtypes = [vp, fp]
    L.odxMlpSetData.argtypes = [vp, fp, fp, i]
    L.odxMlpGradients.argtypes = [vp, fp, fp]
    L.odxMlpTrain.argtypes = [vp, i, i, ctypes.c_float, ctypes.c_uint, fp]
    L.odxMlpPredict.argtypes = [vp, fp, i, fp]
    L.odxMfmaProbe.argtypes = [fp, fp, fp]
    L.odxBatchSetMlp.argtypes = [vp, vp]
    L.odxBatchGetSamples.argtypes = [vp, fp, fp, i]


def mfma_probe(a_lanes, b_lanes):
    L = lib()
    _setup_mlp_api(L)
    a = np.ascontiguousarray(a_lanes, np.float32).reshape(64, 8)
    b = np.ascontiguousarray(b_lanes, np.float32).reshape(64, 8)
    d = np.zeros((64, 4), np.float32)
    _check(L.odxMfmaProbe(_p(a), _p(b), _p(d)), "odxMfmaProbe")
    return d


class Mlp:
    """The learned warm-start network (odxMlp*): bf16 MFMA, fp32 accumulate."""

    def __init__(self, seed=1):
        L = lib()
        _setup_mlp_api(L)
        self.h = L.odxMlpCreate(seed)
        _check(self.h, "odxMlpCreate")
        self.norm = np.concatenate([np.zeros(16), np.ones(16), np.zeros(3), np.ones(3)]).astype(np.float32)

    def close(self):
        if getattr(self, "h", None):
            lib().odxMlpDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_normalization(self, norm38):
        self.norm = np.ascontiguousarray(norm38, np.float32)
        _check(lib().odxMlpSetNormalization(self.h, _p(self.norm)), "odxMlpSetNormalization")

    def set_gate(self, gate):
        """1: matched contacts keep their previous impulses (the network predicts
        new contacts only).  Takes effect for batches attached afterwards."""
        _check(lib().odxMlpSetGate(self.h, int(gate)), "odxMlpSetGate")

    def params(self):
        out = np.zeros(NPARAM, np.float32)
        _check(lib().odxMlpGetParams(self.h, _p(out)), "odxMlpGetParams")
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        _check(lib().odxMlpSetParams(self.h, _p(a)), "odxMlpSetParams")

    def set_data(self, X, T):
        self._X = np.ascontiguousarray(X, np.float32)
        self._T = np.ascontiguousarray(T, np.float32)
        _check(lib().odxMlpSetData(self.h, _p(self._X), _p(self._T), len(self._X)), "odxMlpSetData")

    def gradients(self):
        g = np.zeros(NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().odxMlpGradients(self.h, _p(g), _p(loss)), "odxMlpGradients")
        return g, float(loss[0])

    def train(self, steps, batch=4096, lr=1e-3, seed=7):
        losses = np.zeros(steps, np.float32)
        _check(lib().odxMlpTrain(self.h, steps, batch, float(lr), seed, _p(losses)), "odxMlpTrain")
        return losses

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), 3), np.float32)
        _check(lib().odxMlpPredict(self.h, _p(X), len(X), _p(Y)), "odxMlpPredict")
        return Y


def _batch_set_mlp(self, mlp):
    L = lib()
    _setup_mlp_api(L)
    self._mlp = mlp
    _check(L.odxBatchSetMlp(self.h, mlp.h if mlp is not None else None), "odxBatchSetMlp")


def _batch_samples(self, cap=None):
    L = lib()
    _setup_mlp_api(L)
    cap = cap or int(self.row_start[-1])
    X = np.zeros((cap, FEAT), np.float32)
    T = np.zeros((cap, 3), np.float32)
    n = L.odxBatchGetSamples(self.h, _p(X), _p(T), cap)
    return X[:n], T[:n]


Batch.set_mlp = _batch_set_mlp
Batch.samples = _batch_samples
