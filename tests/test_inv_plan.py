"""The plan of side tasks that bta_potrf_inverse runs inside the Cholesky's step
launches (botorch_amd/csrc/inv_plan.h), replayed on the CPU in FP64 block by
block: every operand a task reads must be final before its step starts (L
block column k after step k, X's diagonal block r after step r, X_rc after
its finish task's step), no two pieces of work of one step may touch a block
that one of them writes (the step's strip, trailing update and side tasks all
run at once), every tile's partial chunks start with the storing one, and the
result must equal inv(L).  The GPU tests run the same plan through the
kernels (tests/test_gpu_kernels.py)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GXX = shutil.which("g++")
LEAF = 64


@pytest.fixture(scope="module")
def ivp(tmp_path_factory):
    if GXX is None:
        pytest.skip("g++ not installed")
    so = tmp_path_factory.mktemp("ivp") / "libivp.so"
    src = os.path.join(HERE, "native", "inv_plan_host.cpp")
    subprocess.run([GXX, "-O2", "-std=c++17", "-fPIC", "-shared", src, "-o", str(so)],
                   check=True, capture_output=True, timeout=300)
    lib = ctypes.CDLL(str(so))
    lib.ivp_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return lib


def plan(lib, nb, ncu):
    nsteps = ctypes.c_int(0)
    n = lib.ivp_size(nb, ncu, ctypes.byref(nsteps))
    assert n >= 0, "plan missed a deadline"
    tasks = np.zeros((max(n, 1), 4), np.int32)
    off = np.zeros(nsteps.value + 1, np.int32)
    units = np.zeros(nsteps.value, np.int32)
    npart = np.zeros(nsteps.value, np.int32)
    P = ctypes.POINTER(ctypes.c_int)
    lib.ivp_fill(nb, ncu, tasks.ctypes.data_as(P), off.ctypes.data_as(P),
                 units.ctypes.data_as(P), npart.ctypes.data_as(P))
    # the kernel runs a step's first npart tasks as partial tasks, the rest as finish tasks
    for j in range(nsteps.value):
        kinds = tasks[off[j]:off[j + 1], 0] & 15
        assert (kinds[:npart[j]] == 0).all() and (kinds[npart[j]:] == 1).all()
    return tasks[:n], off, units


def replay(lib, nb, ncu, b=4, seed=0):
    """Runs the plan on block size b (the schedule does not depend on it)."""
    rng = np.random.default_rng(seed)
    n = nb * b
    M = rng.standard_normal((n, n))
    L = np.linalg.cholesky(M @ M.T / n + np.eye(n))
    blk = lambda A, i, j: A[i * b:(i + 1) * b, j * b:(j + 1) * b]
    X = np.full((n, n), np.nan)
    tasks, off, _ = plan(lib, nb, ncu)
    final_x = {}        # (r, c) -> step after which X_rc is final
    acc_state = {}      # (q, c) -> next leaf expected by the tile's sum
    for j in range(nb + 2):
        reads, writes = [], []
        # the step's own work: strip of column j (L blocks (b, j), X_jj and the
        # zero block right of it), trailing update of L blocks right of column j
        if j < nb:
            writes += [("L", r, j) for r in range(j, nb)]
            writes += [("X", j, j)] + ([("X", j, j + 1)] if j + 1 < nb else [])
            reads += [("L", r, j - 1) for r in range(j, nb)] if j > 0 else []
            if j > 0:
                writes += [("L", r, c) for r in range(j + 1, nb) for c in range(j + 1, r + 1)]
        done_this_step = []
        for t in tasks[off[j]:off[j + 1]]:
            kind, first = t[0] & 15, (t[0] >> 4) & 1
            if kind == 0:
                q, c, k0, k1 = t[1], t[2], t[3] & 0xffff, t[3] >> 16
                assert k1 + 2 <= j <= 2 * q + 1, (j, t)
                assert 0 <= c <= k0 < k1 <= 2 * q - 1
                exp = acc_state.get((q, c), c)
                assert k0 == exp and bool(first) == (k0 == c), (j, t, exp)
                acc_state[(q, c)] = k1
                rows = [r for r in (2 * q, 2 * q + 1) if r < nb]
                for k in range(k0, k1):
                    assert k < j  # L block column k final after step k
                    assert final_x.get((k, c), 10 ** 9) < j or (k == c and k < j), (j, t, k)
                    reads += [("L", r, k) for r in rows] + [("X", k, c)]
                for r in rows:
                    writes.append(("X", r, c))
                    S = sum(blk(L, r, k) @ blk(X, k, c) for k in range(k0, k1))
                    if first:
                        blk(X, r, c)[:] = S
                    else:
                        blk(X, r, c)[:] += S
            else:
                r, c = t[1], t[2]
                q = r // 2
                assert j == r + 2 and 0 <= c < r
                accumulated = q >= 1 and c <= 2 * q - 2
                if accumulated:
                    assert acc_state.get((q, c)) == 2 * q - 1, (j, t)  # all chunks applied
                    S = blk(X, r, c).copy()
                else:
                    S = np.zeros((b, b))
                for k in range(max(c, 2 * q - 1), r):
                    assert k < j and (final_x.get((k, c), 10 ** 9) < j or k == c)
                    reads += [("L", r, k), ("X", k, c)]
                    S = S + blk(L, r, k) @ blk(X, k, c)
                assert r < j
                reads.append(("X", r, r))
                writes.append(("X", r, c))
                blk(X, r, c)[:] = -blk(X, r, r) @ S
                done_this_step.append((r, c))
        # hazards: a block written in this step is touched by nothing else in it
        wset = {}
        for w in writes:
            wset[w] = wset.get(w, 0) + 1
        assert max(wset.values(), default=1) == 1, "two writers of one block in a step"
        assert not (set(reads) & set(wset)), (j, sorted(set(reads) & set(wset))[:4])
        if j < nb:
            blk(X, j, j)[:] = np.linalg.inv(blk(L, j, j))
            final_x[(j, j)] = j
        for rc in done_this_step:
            final_x[rc] = j
    for r in range(nb):
        for c in range(r):
            assert (r, c) in final_x, (r, c)
    Xt = np.tril(X)
    ref = np.linalg.inv(L)
    err = np.abs(Xt - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 6, 7, 8, 9, 16, 17, 33, 64])
@pytest.mark.parametrize("ncu", [256, 32])
def test_plan_replay(ivp, nb, ncu):
    replay(ivp, nb, ncu)


@pytest.mark.parametrize("nb", [64, 128, 256])
def test_plan_every_deadline_met(ivp, nb):
    tasks, off, units = plan(ivp, nb, 256)
    # each (row pair, column) tile: leaves [c, 2q - 1) covered exactly once
    cover = {}
    for t in tasks:
        if t[0] & 15 == 0:
            q, c, k0, k1 = t[1], t[2], t[3] & 0xffff, t[3] >> 16
            cover.setdefault((q, c), []).append((k0, k1))
    for q in range(1, (nb + 1) // 2):
        for c in range(0, 2 * q - 1):
            segs = sorted(cover[(q, c)])
            assert segs[0][0] == c and segs[-1][1] == 2 * q - 1
            assert all(a[1] == b[0] for a, b in zip(segs, segs[1:]))
    nfin = sum(1 for t in tasks if t[0] & 15 == 1)
    assert nfin == nb * (nb - 1) // 2
