"""Batches of independent matrix-ODE solves of uneven sizes and ranks (needs an MI355X).

A batch step queues every member's step before it waits for any of them; each
member runs exactly the kernels a context of its own runs, in the same order,
so the results must agree bit for bit with one context per problem, and with
the CPU oracle to the FP64 bound (relative Frobenius error <= 1e-10).
"""
import sys

import numpy as np
import pytest

from oracle import dlra_de
from oracle.data import random_factors
from oracle.metrics import rel_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SIZES = [(48, 48, 3), (130, 97, 6), (64, 64, 8), (256, 256, 4), (77, 120, 5), (300, 300, 16)]


def problems(lri, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k, (n, m, r) in enumerate(SIZES):
        if k % 2 == 0:  # constant coefficients (analytic eigenbasis)
            A = lri.laplacian_1d(n, 40.0, periodic=(k % 4 == 0))
            B = lri.laplacian_1d(m, 40.0)
        else:  # variable coefficients (RK4 column kernel)
            A = lri.SymTridiagonal(-rng.uniform(40, 80, n), rng.uniform(10, 20, n - 1))
            B = lri.SymTridiagonal(-rng.uniform(40, 80, m), rng.uniform(10, 20, m - 1))
        F = lri.LyapunovOperator(A, B, rng.standard_normal((n, 1)), rng.standard_normal((m, 1)))
        u = random_factors(rng, n, m, r, decay=0.5)
        out.append((F, u))
    return out


def run_alone(lri, F, u, steps, dt, N, tol, rmax):
    n, m = u.U.shape[0], u.V.shape[0]
    with lri.native.Context(n, m, min(rmax, n, m)) as c:
        c.set_state(u.U, u.S, u.V)
        c.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic, F.G, F.H)
        ranks = [c.step_de_bug_ra(dt, N, tol, rmax) for _ in range(steps)]
        return c.get_state(), ranks


def test_batch_rank_adaptive_matches_single_contexts(lri):
    probs = problems(lri)
    dt, tol, rmax, steps = 2e-3, 1e-6, 24, 4
    Ns = [lri.rk4_substeps(F.spectral_bound(), dt) for F, _ in probs]
    say(f"case: batch of {len(probs)} RA problems, sizes {SIZES}, {steps} steps")
    with lri.native.Batch([u.U.shape[0] for _, u in probs], [u.V.shape[0] for _, u in probs],
                          [min(rmax, u.U.shape[0], u.V.shape[0]) for _, u in probs]) as b:
        for (F, u), c in zip(probs, b.members):
            c.set_state(u.U, u.S, u.V)
            c.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic, F.G, F.H)
        ranks = [b.step_de_bug_ra(dt, Ns, tol, rmax) for _ in range(steps)]
        got = [c.get_state() for c in b.members]
    for i, (F, u) in enumerate(probs):
        ref, rk = run_alone(lri, F, u, steps, dt, Ns[i], tol, rmax)
        assert [int(r[i]) for r in ranks] == rk
        for a, b_ in zip(got[i], ref):
            assert a.shape == b_.shape and np.array_equal(a, b_), f"problem {i}"
    # one member against the oracle
    F, u = probs[1]
    host = dlra_de.LyapunovOperator(F.A.dense(), F.B.dense(), F.G, F.H)
    U, S, V = u.U, u.S, u.V
    for _ in range(steps):
        U, S, V, _ = dlra_de.rank_adaptive_bug_de_step(host, U, S, V, 0.0, dt, Ns[1], tol, rmax)
    e = rel_err(*got[1], U, S, V)
    say(f"  member 1 vs oracle after {steps} steps: rank {S.shape[0]}, rel err {e:.2e}")
    assert got[1][1].shape == S.shape and e <= 1e-10


def test_solve_batch_bug(lri):
    probs = problems(lri, seed=1)
    dt, tf = 1e-3, 3e-3
    P = [lri.MatrixDEProblem(F, lri.SVDLikeRepresentation(u.U, u.S, u.V), (0.0, tf)) for F, u in probs]
    sols = lri.solve_batch(P, lri.UnconventionalAlgorithm(), dt, substeps=6)
    assert len(sols) == len(P)
    for (F, u), s in zip(probs, sols):
        assert len(s.Y) == 2 and abs(s.t[-1] - tf) < 1e-15
        host = dlra_de.LyapunovOperator(F.A.dense(), F.B.dense(), F.G, F.H)
        U, S, V = u.U, u.S, u.V
        for k in range(3):
            U, S, V = dlra_de.bug_de_step(host, U, S, V, k * dt, dt, 6)
        y = s.Y[-1]
        assert rel_err(y.U, y.S, y.V, U, S, V) <= 1e-10


def test_batch_graph_replays_match_direct_launches(lri):
    """With rmax = r and tol = 0 the rank never changes, so every member replays the
    graph captured at its first step; each replay (with host reads of the state
    in between) must equal the same batch stepped with direct launches."""
    probs = problems(lri, seed=2)
    dt, steps = 2e-3, 6
    Ns = [lri.rk4_substeps(F.spectral_bound(), dt) for F, _ in probs]
    rs = [u.S.shape[0] for _, u in probs]
    say(f"case: graph replays, {len(probs)} members, fixed ranks {rs}, {steps} steps")
    states = {}
    for graphs in (True, False):
        with lri.native.Batch([u.U.shape[0] for _, u in probs], [u.V.shape[0] for _, u in probs], rs) as b:
            b.use_graphs(graphs)
            for (F, u), c in zip(probs, b.members):
                c.set_state(u.U, u.S, u.V)
                c.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic, F.G, F.H)
            seq = []
            for _ in range(steps):
                # rmax = the member's own capacity (= its rank): pass the largest; each context clips it
                ranks = b.step_de_bug_ra(dt, Ns, 0.0, max(rs))
                assert list(ranks) == rs
                seq.append([c.get_state() for c in b.members])
            states[graphs] = seq
    for k in range(steps):
        for i in range(len(probs)):
            for a, b_ in zip(states[True][k][i], states[False][k][i]):
                assert np.array_equal(a, b_), f"step {k}, member {i}"


def cfg3_problems(count, seed):
    """The batched workload's distribution: n = m log-uniform in [256, 4096], initial ranks
    uniform in {4..64} with the first three forced to 5, 8, 12 and the last four to 48, 56, 64,
    64, Y0 = U diag(10^(-i/4)) V'."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(256), np.log(4096), count)).astype(np.int64)
    rs = rng.integers(4, 65, count)
    rs[:3] = [5, 8, 12]
    rs[-4:] = [48, 56, 64, 64]
    ns[-4] = 300  # one wide member small enough for the oracle
    out = []
    for n, r in zip(ns, rs):
        U, _ = np.linalg.qr(rng.standard_normal((n, r)))
        V, _ = np.linalg.qr(rng.standard_normal((n, r)))
        out.append((int(n), int(r), U, np.diag(10.0 ** (-np.arange(r) / 4.0)), V))
    return out


def test_batch_cfg3_distribution(lri):
    """32 members drawn like the batched workload (dX/dt = L X + X L', L the Dirichlet
    Laplacian n^2 tridiag(1, -2, 1), tol 1e-8, dt 1e-4, rmax 64), two rank-adaptive batch
    steps: every member bit for bit equal to a context of its own (the wide members run the
    96 x 96 and 128 x 128 SVDs), and the three members of initial rank 5, 8 and 12 against the
    oracle.  (With the n^2 scale the steps sit at RK4's stability limit, and for initial ranks
    >= 30 the singular values reach 1e-8, where the truncation decides: two evaluations of the
    oracle that differ only in summation order there pick different ranks, so those members
    are compared with their own contexts only.)"""
    import scipy.sparse as sp

    count, dt, tol, rmax, steps = 32, 1e-4, 1e-8, 64, 2
    probs = cfg3_problems(count, seed=3)
    Ns = [lri.rk4_substeps(2.0 * lri.laplacian_1d(n).gershgorin(), dt) for n, *_ in probs]
    say(f"case: cfg3-shaped batch, {count} members, n in [{min(p[0] for p in probs)}, {max(p[0] for p in probs)}], "
        f"ranks {[p[1] for p in probs]}")
    with lri.native.Batch([p[0] for p in probs], [p[0] for p in probs], [rmax] * count) as b:
        for (n, r, U, S, V), c in zip(probs, b.members):
            Lop = lri.laplacian_1d(n)
            c.set_operator_lyapunov(Lop.d, Lop.e, False, Lop.d, Lop.e, False)
            c.set_state(U, S, V)
        ranks = [b.step_de_bug_ra(dt, Ns, tol, rmax) for _ in range(steps)]
        assert not b.member_status().any()
        got = [c.get_state() for c in b.members]
    say(f"  ranks after {steps} steps: {list(ranks[-1])}")
    for i, (n, r, U, S, V) in enumerate(probs):
        with lri.native.Context(n, n, rmax) as c:
            Lop = lri.laplacian_1d(n)
            c.set_operator_lyapunov(Lop.d, Lop.e, False, Lop.d, Lop.e, False)
            c.set_state(U, S, V)
            rk = [c.step_de_bug_ra(dt, Ns[i], tol, rmax) for _ in range(steps)]
            ref = c.get_state()
        assert [int(x[i]) for x in ranks] == rk, f"member {i}"
        for a, b_ in zip(got[i], ref):
            assert a.shape == b_.shape and np.array_equal(a, b_), f"member {i}"
    check = [0, 1, 2]
    for i in check:
        n, r, U, S, V = probs[i]
        L = sp.csr_matrix(sp.diags([np.ones(n - 1), -2.0 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]) * float(n * n))
        op = dlra_de.LowRankLyapunov(L, L, np.zeros((n, 1)), np.zeros((n, 1)))
        # the oracle from the initial basis and from one perturbed at the rounding level: how far
        # the problem itself lets two exact-arithmetic-equivalent evaluations drift apart
        Up, _ = np.linalg.qr(U + 1e-14 * np.random.default_rng(1).standard_normal(U.shape))
        a, b = (U, S, V), (Up, S, V)
        for _ in range(steps):
            a = dlra_de.rank_adaptive_bug_de_step_lowrank(op, *a, 0.0, dt, Ns[i], tol, rmax)[:3]
            b = dlra_de.rank_adaptive_bug_de_step_lowrank(op, *b, 0.0, dt, Ns[i], tol, rmax)[:3]
        assert a[1].shape == b[1].shape
        norm0, norm_o, norm_e = (np.linalg.norm(x) for x in (S, a[1], got[i][1]))
        if norm_o <= 1e-20 * norm0:
            # the projected flow damped the whole state (its Galerkin operator on the random
            # initial bases has only large eigenvalues): nothing left to compare but that
            say(f"  member {i} (n={n}, r0={r}): state damped to |S| = {norm_o:.1e} (oracle), {norm_e:.1e} (engine)")
            assert norm_e <= 1e-20 * norm0
            continue
        sens = rel_err(*a, *b)
        e = rel_err(*got[i], *a) if got[i][1].shape == a[1].shape else float("inf")
        say(f"  member {i} (n={n}, r0={r}) vs oracle: rank {a[1].shape[0]}, rel err {e:.2e} "
            f"(oracle perturbed at 1e-14: {sens:.2e})")
        assert got[i][1].shape == a[1].shape and e <= max(1e-10, 50 * sens)


def resident_batch(lri, probs, rcaps, skip=(), scale=None):
    b = lri.native.Batch([p[0] for p in probs], [p[0] for p in probs], rcaps)
    for k, ((n, r, U, S, V), c) in enumerate(zip(probs, b.members)):
        Lop = lri.laplacian_1d(n) if scale is None else lri.laplacian_1d(n, scale)
        c.set_operator_lyapunov(Lop.d, Lop.e, False, Lop.d, Lop.e, False)
        c.set_resident_eigenbasis(True)
        if k not in skip:
            c.set_state(U, S, V)
    return b


def test_merged_batch_matches_single_contexts(lri):
    """Merged launches (every member in its resident eigenbasis): 40 members drawn like the
    batched workload, three rank-adaptive steps at tol 1e-8 (ranks change), one member left
    without a state (it reports LRI_EINVAL and the others advance).  Every other member
    equals a resident context of its own bit for bit, and the step is a few hundred
    launches for all members, not one sequence per member."""
    count, dt, tol, rmax, steps, bad = 40, 1e-4, 1e-8, 64, 3, 7
    probs = cfg3_problems(count, seed=5)
    Ns = [lri.rk4_substeps(2.0 * lri.laplacian_1d(n).gershgorin(), dt) for n, *_ in probs]
    say(f"case: merged batch, {count} members, ranks {[p[1] for p in probs]}, member {bad} without a state")
    with resident_batch(lri, probs, [rmax] * count, skip=(bad,)) as b:
        ranks = []
        for _ in range(steps):
            with pytest.raises(ValueError, match=f"problem {bad}: state not set"):
                b.step_de_bug_ra(dt, Ns, tol, rmax)
            st = b.member_status()
            assert st[bad] == lri.native.EINVAL and not np.delete(st, bad).any()
            ranks.append([c.rank() for c in b.members])
        merged, launches = b.launch_stats()
        got = [c.get_state() if k != bad else None for k, c in enumerate(b.members)]
    say(f"  {merged} merged steps, {launches} launches in the last; ranks {ranks[-1]}")
    assert merged == steps and 0 < launches < 8 * count
    for i, (n, r, U, S, V) in enumerate(probs):
        if i == bad:
            continue
        with lri.native.Context(n, n, rmax) as c:
            Lop = lri.laplacian_1d(n)
            c.set_operator_lyapunov(Lop.d, Lop.e, False, Lop.d, Lop.e, False)
            c.set_resident_eigenbasis(True)
            c.set_state(U, S, V)
            rk = [c.step_de_bug_ra(dt, Ns[i], tol, rmax) for _ in range(steps)]
            ref = c.get_state()
        assert [x[i] for x in ranks] == rk, f"member {i}"
        for a, b_ in zip(got[i], ref):
            assert a.shape == b_.shape and np.array_equal(a, b_), f"member {i}"


def test_merged_batch_graph_replays(lri):
    """tol = 0 with each member's capacity at its rank, and a mild operator (the Laplacian
    without the n^2 scale, so that no mode decays to zero): the ranks stay, the merged list
    repeats and is replayed as a captured graph from the second step on.  Six steps with the
    states read back between them must equal the same batch with direct launches, and the
    merged batch must equal the batch without merged launches (one stream per member)."""
    count, dt, steps = 24, 1e-4, 6
    probs = cfg3_problems(count, seed=6)
    rs = [p[1] for p in probs]
    Ns = [lri.rk4_substeps(2.0 * lri.laplacian_1d(n, 1.0).gershgorin(), dt) for n, *_ in probs]
    say(f"case: merged graph replays, {count} members, fixed ranks {rs}, {steps} steps")
    seqs, bad = {}, []
    for mode in ("direct", "graphs", "unmerged"):
        with resident_batch(lri, probs, rs, scale=1.0) as b:
            b.use_graphs(mode != "direct")
            b.use_merged_launches(mode != "unmerged")
            seq = []
            for k in range(steps):
                got = [int(x) for x in b.step_de_bug_ra(dt, Ns, 0.0, 64)]
                wrong = [i for i in range(count) if got[i] != rs[i]]
                say(f"  {mode} step {k}: members with another rank: {[(i, rs[i], got[i]) for i in wrong]}")
                bad += [(mode, k, i) for i in wrong]
                seq.append([c.get_state() for c in b.members])
            merged, _ = b.launch_stats()
            assert merged == (0 if mode == "unmerged" else steps)
            seqs[mode] = seq
    assert not bad, bad
    for mode in ("graphs", "unmerged"):
        for k in range(steps):
            for i in range(count):
                for a, b_ in zip(seqs["direct"][k][i], seqs[mode][k][i]):
                    assert np.array_equal(a, b_), f"{mode}: step {k}, member {i}"
