"""The candidate-restricted iteration of the learned proposal, through the
C-ABI:
  - with every coarse cell as a candidate it reproduces the dense FP16
    iteration (same contractions, summation order aside);
  - with top-k candidates, the audit subset and whole-table fallback it
    matches the FP64 oracle restricted to the same rotations (1e-4 absolute
    on responsibilities, 1e-4 relative on scaling and model);
  - the search mode follows the audited captured mass, over two iterations;
  - a network trained on EMC's own responsibilities captures far more of the
    posterior in its top-k cells than the untrained network;
  - when more of the audit than audit_max_fail misses min_mass, the iteration
    runs on the whole table (identical to the dense FP16 iteration), and the
    next candidates are re-centred on that whole-table posterior;
  - with shared cells (the cells holding the patterns' last posteriors),
    every pattern is also scored on the shared cells' rotations and the
    iteration matches the FP64 oracle restricted to candidates + shared
    rotations (+ the whole table for the audited patterns);
  - scoring a group's common candidate cells as one dense block gives the
    same iteration, bit for bit, as scoring every candidate cell alone;
  - clearing only the previous iteration's nonzero responsibilities before
    writing the new ones leaves the same matrix, bit for bit, as rewriting
    all of it, over consecutive iterations.
"""
import ctypes as C

import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

pytestmark = pytest.mark.gpu

SEED = 11


@pytest.fixture(scope="module")
def prob():
    # 32^3 model, 32x32 patterns, n=4 fine table (3,240), truth-model start
    return problems.make(side=32, n_patterns=300, n_rot=4, n_blobs=20, radius=6.0, photons=2000.0,
                         phi_sigma=0.2, beamstop=3.0, seed=3, model="truth")


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    opts(proposal="off")
    yield
    opts(proposal="off")
    E.lib().emc_hip_clear_error()


def opts(precision="fp16", insert="deterministic", proposal="off", audit="0", min_mass="0.999", gate="audit",
         track="off", max_fail="1", mstep="auto", shared="0", group_min="0", fill="delta", gemm_tail="split"):
    E.set_option("precision", precision)
    E.set_option("insert", insert)
    E.set_option("interp", "nearest")
    E.set_option("proposal", proposal)
    E.set_option("audit_fraction", audit)
    E.set_option("min_mass", min_mass)
    E.set_option("proposal_gate", gate)
    E.set_option("proposal_track", track)
    E.set_option("audit_max_fail", max_fail)
    E.set_option("mstep", mstep)
    E.set_option("proposal_shared", shared)
    E.set_option("proposal_group_min", group_min)
    E.set_option("respons_fill", fill)
    E.set_option("gemm_tail", gemm_tail)


def init_cells(P, topk, n_coarse=2, batch=64):
    L = E.lib()
    assert L.emc_hip_proposal_init(E.fptr(P.q), P.R, n_coarse, P.side, topk, batch, SEED) == 0
    E.check()
    nc = L.emc_hip_proposal_cells(None, None)
    off = np.zeros(nc + 1, np.int32)
    rot = np.zeros(P.R, np.int32)
    assert L.emc_hip_proposal_cells(E.iptr(off), E.iptr(rot)) == nc
    return nc, off, rot


def stats():
    out = (C.c_double * 22)()
    assert E.lib().emc_hip_proposal_stats(out, 22) == 22
    return list(out)


def audited(N, iteration, frac):
    L = E.lib()
    return np.array([i for i in range(N) if L.emc_uniform(SEED, 0xA0D17, iteration, i) < frac], np.int64)


def run(P, model=None, scaling=None, iterations=1):
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model if model is None else model,
                        scaling=scaling)
    out = []
    for _ in range(iterations):
        it = prob.iteration(0.4)
        out.append({"P": prob.respons(), "scaling": prob.scaling(), "model": prob.model(), "it": it,
                    "stats": stats()})
    prob.free()
    return out


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def cell_mass(Pm, off, rot):
    """[C][N] posterior mass per coarse cell."""
    return np.stack([Pm[rot[off[c]:off[c + 1]]].sum(axis=0) for c in range(len(off) - 1)])


def test_all_cells_reproduce_dense(prob):
    """Same operands and the same k order in both kernels (the dense GEMM
    unsplit along k): the candidate iteration equals the dense one to FP32
    rounding of the epilogues."""
    P = prob
    opts(gemm_tail="whole")
    dense = run(P)[0]
    nc, off, rot = init_cells(P, topk=8)
    cand = np.tile(np.arange(nc, dtype=np.int32), (P.N, 1))
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, nc) == 0
    opts(proposal="on", gemm_tail="whole")
    cand_run = run(P)[0]
    st = cand_run["stats"]
    assert st[0] == 1 and st[11] == 0 and st[6] == P.N * P.R
    dP = np.abs(cand_run["P"] - dense["P"]).max()
    ds = rel(cand_run["scaling"], dense["scaling"])
    dm = rel(cand_run["model"], dense["model"])
    dl = abs(cand_run["it"]["likelihood"] - dense["it"]["likelihood"]) / abs(dense["it"]["likelihood"])
    print(f"all cells vs dense FP16: P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}, sum PlogP {dl:.2e}")
    assert dP < 1e-6 and ds < 1e-6 and dm < 1e-5 and dl < 1e-6


def oracle_candidates(P, off, rot, topk):
    ref = O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                          P.x, P.y, P.z)
    cm = cell_mass(ref["P"], off, rot)
    return np.argsort(-cm, axis=0, kind="stable")[:topk].T.astype(np.int32).copy(), ref


@pytest.mark.parametrize("audit,mstep", [("0", "dense"), ("0.3", "dense"), ("0.3", "sparse")])
def test_topk_with_audit_matches_oracle(prob, audit, mstep):
    """mstep "dense": the candidate M-step; "sparse": the M-step over the
    nonzero responsibilities of the scattered candidate + subset posterior."""
    P = prob
    topk = 6
    nc, off, rot = init_cells(P, topk)
    cand, ref_dense = oracle_candidates(P, off, rot, topk)
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
    opts(proposal="on", audit=audit, mstep=mstep)
    g = run(P)[0]
    sub = audited(P.N, 1, float(audit))
    allowed = O.allowed_from_candidates(cand, off, rot, P.R, sub)
    ref = O.o64_proposal_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                                   P.x, P.y, P.z, allowed)
    st = g["stats"]
    dP = np.abs(g["P"] - ref["P"]).max()
    ds = rel(g["scaling"], ref["scaling"])
    dm = rel(g["model"], ref["model"])
    assert E.lib().emc_hip_mstep_nnz() == (int((g["P"] != 0).sum()) if mstep == "sparse" else -1)
    print(f"top-{topk}, audit {audit}, mstep {mstep}: {len(sub)} audited, pairs {st[6]:.0f} of {P.N * P.R}; "
          f"P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}; captured mass min {st[4]:.6f} mean {st[5]:.6f}")
    assert st[2] == len(sub) and st[11] == len(sub)
    assert st[6] == allowed.sum()
    assert dP < 1e-4 and ds < 1e-4 and dm < 1e-4
    # search modes: audited patterns whose candidates hold < 0.999 of the
    # whole-table posterior go to full search
    mode = np.zeros(P.N, np.int32)
    assert E.lib().emc_hip_proposal_get(None, E.iptr(mode), None, P.N) == 0
    mass = np.array([ref_dense["P"][allowed_from(cand[i], off, rot), i].sum() for i in range(P.N)])
    clear = np.abs(mass - 0.999) > 1e-4
    expect = np.zeros(P.N, np.int32)
    expect[sub] = (mass[sub] < 0.999)
    assert np.array_equal(mode[clear], expect[clear])
    assert st[3] == expect.sum() or not clear[sub].all()


def allowed_from(cells, off, rot):
    return np.concatenate([rot[off[c]:off[c + 1]] for c in cells])


def test_fallback_over_two_iterations(prob):
    P = prob
    topk = 3
    nc, off, rot = init_cells(P, topk)
    # deliberately poor candidates: cells ranked 2..4 by posterior mass
    cm = cell_mass(O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                                   P.x, P.y, P.z)["P"], off, rot)
    cand = np.argsort(-cm, axis=0, kind="stable")[2:2 + topk].T.astype(np.int32).copy()
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
    opts(proposal="on", audit="0.5")  # fixed candidates: the fallback alone
    g1, g2 = run(P, iterations=2)
    mode1 = np.zeros(P.N, np.int32)
    # modes after iteration 2; reconstruct those after iteration 1 from the
    # oracle to predict the subset of iteration 2
    sub1 = audited(P.N, 1, 0.5)
    allowed1 = O.allowed_from_candidates(cand, off, rot, P.R, sub1)
    ref1 = O.o64_proposal_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                                    P.x, P.y, P.z, allowed1)
    full1 = np.zeros(P.N, bool)
    for i in sub1:
        full1[i] = ref1["P"][allowed_from(cand[i], off, rot), i].sum() < 0.999
    print(f"iteration 1: {len(sub1)} audited, {full1.sum()} to full search (GPU {g1['stats'][3]:.0f})")
    assert g1["stats"][3] == full1.sum()
    assert g2["stats"][1] == full1.sum()  # whole-table search in iteration 2
    sub2 = np.union1d(audited(P.N, 2, 0.5), np.nonzero(full1)[0])
    assert g2["stats"][11] == len(sub2)
    # iteration 2 from the GPU's own iteration-1 model and scaling
    allowed2 = O.allowed_from_candidates(cand, off, rot, P.R, sub2)
    ref2 = O.o64_proposal_iteration(g1["model"].astype(np.float64), P.images, P.mask, g1["scaling"], P.side, P.q,
                                    P.w, P.x, P.y, P.z, allowed2)
    dP = np.abs(g2["P"] - ref2["P"]).max()
    dm = rel(g2["model"], ref2["model"])
    print(f"iteration 2: {len(sub2)} on the whole table; P {dP:.2e}, model {dm:.2e}")
    assert dP < 1e-4 and dm < 1e-4
    del mode1


def test_trained_network_gain(prob):
    P = prob
    topk = 8
    nc, off, rot = init_cells(P, topk, batch=64)
    L = E.lib()
    opts()
    p = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    p.get_slices()
    p.likelihood(0.4)
    p.normalize_respons()
    Pm = p.respons().astype(np.float64)
    cm = cell_mass(Pm, off, rot)

    def captured():
        cand = np.zeros((P.N, topk), np.int32)
        pm = np.zeros(P.N, np.float32)
        assert L.emc_hip_proposal_predict(p.d_images.vp, p.d_mask.vp, P.side ** 2, P.N) == 0
        assert L.emc_hip_proposal_get(E.iptr(cand), None, E.fptr(pm), P.N) == 0
        return np.take_along_axis(cm, cand.T, axis=0).sum(axis=0), pm

    m0, pm0 = captured()
    loss = C.c_double()
    losses = []
    for _ in range(16):
        assert L.emc_hip_proposal_train(p.d_images.vp, p.d_mask.vp, P.side ** 2, p.d_respons.vp, P.N, P.R, 5, 2e-3,
                                        C.byref(loss)) == 0
        losses.append(loss.value)
    m1, pm1 = captured()
    p.free()
    E.check()
    print(f"training loss {losses[0]:.3f} -> {losses[-1]:.3f}; captured mass in top-{topk} of {nc} cells: "
          f"untrained mean {m0.mean():.3f} (>=0.999: {np.mean(m0 >= 0.999):.3f}), "
          f"trained mean {m1.mean():.3f} (>=0.999: {np.mean(m1 >= 0.999):.3f}); "
          f"predicted top-k mass {pm0.mean():.3f} -> {pm1.mean():.3f}")
    assert losses[-1] < 0.5 * losses[0]
    assert m1.mean() > 0.9 and m1.mean() > m0.mean() + 0.5
    assert np.mean(m1 >= 0.99) > np.mean(m0 >= 0.99) + 0.5


def test_prediction_is_topk_of_network_output(prob):
    """Candidates = the network's k largest outputs (descending, ties to the
    lower cell), mass = softmax probability they hold."""
    P = prob
    topk = 8
    nc, off, rot = init_cells(P, topk, batch=64)
    L = E.lib()
    d_img = E.DeviceArray.from_host(P.images.reshape(-1))
    d_mask = E.DeviceArray.from_host(P.mask)
    n = 40
    assert L.emc_hip_proposal_predict(d_img.vp, d_mask.vp, P.side ** 2, P.N) == 0
    cand = np.zeros((P.N, topk), np.int32)
    mass = np.zeros(P.N, np.float32)
    assert L.emc_hip_proposal_get(E.iptr(cand), None, E.fptr(mass), P.N) == 0
    idx = np.arange(5, 5 + n, dtype=np.int32)
    logits = np.zeros((nc, n), np.float32)
    assert L.emc_hip_cnn_forward(d_img.vp, d_mask.vp, P.side ** 2, E.iptr(idx), n, E.fptr(logits)) == 0
    d_img.free()
    d_mask.free()
    E.check()
    ref = np.argsort(-logits, axis=0, kind="stable")[:topk].T
    assert np.array_equal(cand[idx], ref)
    lg = logits.astype(np.float64)
    p = np.exp(lg - lg.max(axis=0))
    p /= p.sum(axis=0)
    ref_mass = np.take_along_axis(p, ref.T, axis=0).sum(axis=0)
    assert np.allclose(mass[idx], ref_mass, rtol=1e-5, atol=1e-6)


def test_tracking_recentres_on_the_posterior(prob):
    """With deliberately poor candidates (cells ranked 3..8 by posterior mass)
    and no audit, the first proposal iteration misses most of the mass; the
    candidates of the second are re-centred on the cells that hold the mass
    of the first iteration's posterior (whole-table posterior for the patterns
    scored there), so they capture far more of the next whole-table posterior."""
    P = prob
    topk = 6
    nc, off, rot = init_cells(P, topk)
    ref = O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                          P.x, P.y, P.z)
    cm = cell_mass(ref["P"], off, rot)
    order = np.argsort(-cm, axis=0, kind="stable")
    cand = order[2:2 + topk].T.astype(np.int32).copy()
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
    opts(proposal="on", audit="1", track="on")  # every pattern scored on the whole table
    g1, g2 = run(P, iterations=2)
    new = np.zeros((P.N, topk), np.int32)
    assert E.lib().emc_hip_proposal_get(E.iptr(new), None, None, P.N) == 0
    st2 = g2["stats"]
    # the re-centred candidates hold each pattern's best cell of iteration 1
    # (whole-table posterior, computed here from the GPU's iteration-1 P)
    cm1 = cell_mass(g1["P"].astype(np.float64), off, rot)
    best1 = cm1.argmax(axis=0)
    hit = np.mean([best1[i] in new[i] for i in range(P.N)])
    old_mass = np.take_along_axis(cm1, cand.T, axis=0).sum(axis=0)
    new_mass = np.take_along_axis(cm1, new.T, axis=0).sum(axis=0)
    print(f"re-centred {st2[13]:.0f} patterns; best cell kept {hit:.3f}; iteration-1 mass in old candidates "
          f"{old_mass.mean():.3f}, in re-centred {new_mass.mean():.3f}")
    assert st2[13] > 0.5 * P.N
    assert hit == 1.0
    assert new_mass.mean() > old_mass.mean() + 0.3


def test_audit_sends_iteration_to_whole_table(prob):
    """Poor candidates (cells ranked 3..8): far more than audit_max_fail of
    the audited patterns miss min_mass, so the iteration runs on the whole
    table and equals the dense FP16 iteration; the failing audited patterns
    go to full search; the next iteration's candidates are re-centred on the
    whole-table posterior and hold every pattern's best cell of it."""
    P = prob
    topk = 6
    opts()
    dense = run(P)[0]
    nc, off, rot = init_cells(P, topk)
    cm = cell_mass(O.o64_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                                   P.x, P.y, P.z)["P"], off, rot)
    cand = np.argsort(-cm, axis=0, kind="stable")[2:2 + topk].T.astype(np.int32).copy()
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
    opts(proposal="on", audit="0.2", track="on", max_fail="0.005")
    g1, g2 = run(P, iterations=2)
    st1 = g1["stats"]
    sub = audited(P.N, 1, 0.2)
    print(f"iteration 1: {st1[2]:.0f} audited, {st1[3]:.0f} below min mass, whole table {st1[14]:.0f}")
    assert st1[2] == len(sub) and st1[3] > 0.005 * st1[2]
    assert st1[14] == 1 and st1[6] == P.N * P.R
    dP = np.abs(g1["P"] - dense["P"]).max()
    ds = rel(g1["scaling"], dense["scaling"])
    dm = rel(g1["model"], dense["model"])
    print(f"whole-table iteration vs dense FP16: P {dP:.2e}, scaling {ds:.2e}, model {dm:.2e}")
    assert dP == 0.0 and ds == 0.0 and dm == 0.0
    assert g2["stats"][1] == st1[3]  # the failing audited patterns search the whole table
    new = np.zeros((P.N, topk), np.int32)
    assert E.lib().emc_hip_proposal_get(E.iptr(new), None, None, P.N) == 0
    best1 = cell_mass(g1["P"].astype(np.float64), off, rot).argmax(axis=0)
    hit = np.mean([best1[i] in new[i] for i in range(P.N)])
    print(f"iteration 2: re-centred {g2['stats'][13]:.0f}; best cell of iteration 1 kept {hit:.3f}")
    assert g2["stats"][13] > 0.5 * P.N and hit == 1.0


@pytest.mark.parametrize("mstep", ["dense", "sparse"])
def test_shared_cells_match_oracle(prob, mstep):
    """Iteration 1 scores every pattern on the whole table (audit 1); before
    iteration 2 the cells holding the patterns' best cells become shared and
    the candidates are re-centred around them (excluding them).  Iteration 2
    (30% audit) must equal the FP64 oracle restricted to each pattern's
    candidate cells + all shared rotations (whole table for the audited),
    from the GPU's own iteration-1 model and scaling."""
    P = prob
    topk = 6
    nc, off, rot = init_cells(P, topk)
    cand, _ = oracle_candidates(P, off, rot, topk)
    assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
    opts(proposal="on", audit="1", min_mass="0", track="on", shared="1000", mstep=mstep)
    prob_ = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    prob_.iteration(0.4)
    g1 = {"P": prob_.respons(), "model": prob_.model(), "scaling": prob_.scaling()}
    E.set_option("audit_fraction", "0.3")
    it2 = prob_.iteration(0.4)
    g2 = {"P": prob_.respons(), "model": prob_.model(), "scaling": prob_.scaling(), "stats": stats()}
    prob_.free()
    E.check()
    L = E.lib()
    n_sh = L.emc_hip_proposal_shared(None, 0)
    shared = np.zeros(max(n_sh, 1), np.int32)
    assert L.emc_hip_proposal_shared(E.iptr(shared), n_sh) == n_sh
    shared = shared[:n_sh]
    new = np.zeros((P.N, topk), np.int32)
    assert L.emc_hip_proposal_get(E.iptr(new), None, None, P.N) == 0
    best1 = cell_mass(g1["P"].astype(np.float64), off, rot).argmax(axis=0)
    st = g2["stats"]
    print(f"{n_sh} shared cells of {nc}; iteration-1 best cells {len(np.unique(best1))}; stats shared {st[15]:.0f}")
    assert 0 < n_sh < nc and st[15] == n_sh
    assert set(np.unique(best1)) <= set(shared.tolist())
    assert not np.isin(new, shared).any()  # candidates exclude the shared cells
    assert all(len(np.unique(new[i])) == topk for i in range(P.N))
    sub2 = audited(P.N, 2, 0.3)
    allowed = O.allowed_from_candidates(new, off, rot, P.R, sub2)
    sh_rot = np.concatenate([rot[off[c]:off[c + 1]] for c in shared])
    allowed[sh_rot, :] = True
    ref = O.o64_proposal_iteration(g1["model"].astype(np.float64), P.images, P.mask, g1["scaling"], P.side, P.q,
                                   P.w, P.x, P.y, P.z, allowed)
    dP = np.abs(g2["P"] - ref["P"]).max()
    ds = rel(g2["scaling"], ref["scaling"])
    dm = rel(g2["model"], ref["model"])
    dl = abs(it2["likelihood"] - ref["total"]) / abs(ref["total"])
    print(f"shared, mstep {mstep}: pairs {st[6]:.0f} (allowed {allowed.sum()}); P {dP:.2e}, scaling {ds:.2e}, "
          f"model {dm:.2e}, sum PlogP {dl:.2e}")
    assert st[11] == len(sub2) and st[6] == allowed.sum()
    assert dP < 1e-4 and ds < 1e-4 and dm < 1e-4 and dl < 1e-5


def test_group_blocks_match_per_cell(prob):
    """Every pattern's candidates start with the same three cells (one group
    of 300 patterns, those cells its core), then its own best cells.  The
    iteration with the group scored as a dense block equals the one with
    every candidate cell scored by the per-cell kernel, bit for bit (same
    products, same k order), and matches the restricted FP64 oracle."""
    P = prob
    topk = 6
    nc, off, rot = init_cells(P, topk)
    best, _ = oracle_candidates(P, off, rot, 12)
    common = np.array([0, 7, 19], np.int32)
    cand = np.zeros((P.N, topk), np.int32)
    for i in range(P.N):
        own = [c for c in best[i] if c not in common][:topk - 3]
        cand[i] = np.concatenate([common, own])
    out = {}
    for gmin in ("0", "16"):
        opts(proposal="on", audit="0.3", group_min=gmin)
        assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
        out[gmin] = run(P)[0]
    g0, g1 = out["0"], out["16"]
    print(f"group-block pairs {g1['stats'][17]:.0f} (none without groups: {g0['stats'][17]:.0f})")
    core_rot = sum(off[c + 1] - off[c] for c in common)
    assert g0["stats"][17] == 0 and g1["stats"][17] >= P.N * core_rot
    for k in ("P", "scaling", "model"):
        assert np.array_equal(g0[k], g1[k]), k
    sub = audited(P.N, 1, 0.3)
    allowed = O.allowed_from_candidates(cand, off, rot, P.R, sub)
    ref = O.o64_proposal_iteration(P.model.astype(np.float64), P.images, P.mask, np.ones(P.N), P.side, P.q, P.w,
                                   P.x, P.y, P.z, allowed)
    assert np.abs(g1["P"] - ref["P"]).max() < 1e-4 and rel(g1["model"], ref["model"]) < 1e-4


def test_group_block_core_beyond_one_rotation_tile(prob):
    """A group whose core holds more than 256 rotations (the 28 largest
    cells shared by all 300 patterns, then each pattern's own best cells):
    the group block runs a second 256-rotation tile, and the iteration equals
    the per-cell kernel's bit for bit."""
    P = prob
    topk = 32
    nc, off, rot = init_cells(P, topk)
    sizes = np.diff(off)
    common = np.sort(np.argsort(-sizes, kind="stable")[:28]).astype(np.int32)
    core_rot = int(sizes[common].sum())
    assert core_rot > 256, core_rot
    best, _ = oracle_candidates(P, off, rot, topk + 28)
    cand = np.zeros((P.N, topk), np.int32)
    for i in range(P.N):
        own = [c for c in best[i] if c not in common][:topk - 28]
        cand[i] = np.concatenate([common, own])
    out = {}
    for gmin in ("0", "16"):
        opts(proposal="on", audit="0.3", group_min=gmin)
        assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
        out[gmin] = run(P)[0]
    g0, g1 = out["0"], out["16"]
    print(f"core rotations {core_rot}; group-block pairs {g1['stats'][17]:.0f}")
    assert g0["stats"][17] == 0 and g1["stats"][17] >= P.N * core_rot
    for k in ("P", "scaling", "model"):
        assert np.array_equal(g0[k], g1[k]), k


def test_delta_fill_matches_full(prob):
    """Four proposal iterations (whole-table audit first, then 30%, shared
    cells, re-centring, M-step lists from the posterior's sources): with
    "respons_fill" "delta" iterations 2-4 clear only the previous iteration's
    nonzeros (stat 20 < N * R), and every iteration's responsibilities,
    scaling and model equal those of the whole-matrix rewrite bit for bit."""
    P = prob
    topk = 6
    nc, off, rot = init_cells(P, topk)
    cand, _ = oracle_candidates(P, off, rot, topk)
    out = {}
    for fill in ("full", "delta"):
        assert E.lib().emc_hip_proposal_set_candidates(E.iptr(cand), P.N, topk) == 0
        opts(proposal="on", audit="1", min_mass="0", track="on", shared="1000", mstep="sparse", fill=fill)
        prob_ = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
        res = []
        for it in range(4):
            if it == 1:
                E.set_option("audit_fraction", "0.3")
            prob_.iteration(0.4)
            res.append({"P": prob_.respons(), "scaling": prob_.scaling(), "model": prob_.model(),
                        "cleared": stats()[20], "nnz": E.lib().emc_hip_mstep_nnz()})
        prob_.free()
        E.check()
        out[fill] = res
    for it, (f, d) in enumerate(zip(out["full"], out["delta"])):
        print(f"iteration {it + 1}: cleared full {f['cleared']:.0f} delta {d['cleared']:.0f}, "
              f"M-step nonzeros {d['nnz']}, P nonzeros {(d['P'] != 0).sum()}")
        assert f["cleared"] == P.N * P.R
        if it == 0:
            assert d["cleared"] == P.N * P.R
        else:
            assert d["cleared"] == out["delta"][it - 1]["nnz"] < P.N * P.R
        assert d["nnz"] == (d["P"] != 0).sum()
        for k in ("P", "scaling", "model"):
            assert np.array_equal(f[k], d[k]), (it, k)


def test_headline_path_poisson_matches_oracle():
    """The benchmark's learned path end to end at 32^3: Poisson likelihood,
    FP16 hi/lo, one dense iteration whose responsibilities train the network,
    prediction of the top-k cells, then proposal iterations with posterior
    tracking, shared cells, group blocks, the M-step over nonzero lists, a
    whole-table audit of every pattern in the first proposal iteration and
    of 2% afterwards (fallback of the iteration at 0.5% failures, per-pattern
    full search below 0.999).  Every proposal iteration, from the GPU's own
    model and scaling, must equal the FP64 oracle restricted to the rotations
    it scored (each pattern's candidate cells + the shared cells' rotations,
    the whole table for the audited and full-search patterns, or everything
    when the audit sent the iteration to the whole table)."""
    P = problems.make(side=32, n_patterns=400, n_rot=4, n_blobs=20, radius=6.0, photons=2000.0, phi_sigma=0.3,
                      beamstop=3.0, gap_cross=1, seed=21, model="random")
    topk = 6
    nc, off, rot = init_cells(P, topk, batch=64)
    L = E.lib()
    opts(insert="atomic", mstep="auto", group_min="16", shared="1000", track="on", min_mass="0.999",
         max_fail="0.005", audit="1")
    E.set_option("likelihood", "poisson")
    try:
        prob_ = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
        prob_.iteration(0.4)
        loss = C.c_double()
        assert L.emc_hip_proposal_train(prob_.d_images.vp, prob_.d_mask.vp, P.side ** 2, prob_.d_respons.vp, P.N,
                                        P.R, 4, 3e-3, C.byref(loss)) == 0
        assert L.emc_hip_proposal_predict(prob_.d_images.vp, prob_.d_mask.vp, P.side ** 2, P.N) == 0
        E.set_option("proposal", "on")
        state = {"model": prob_.model(), "scaling": prob_.scaling()}
        mode_prev = np.zeros(P.N, np.int32)
        worst = [0.0, 0.0, 0.0]
        for it in range(1, 5):
            frac = 1.0 if it == 1 else 0.02
            E.set_option("audit_fraction", str(frac))
            res = prob_.iteration(0.4)
            g = {"P": prob_.respons(), "scaling": prob_.scaling(), "model": prob_.model()}
            st = stats()
            cand = np.zeros((P.N, topk), np.int32)
            mode = np.zeros(P.N, np.int32)
            assert L.emc_hip_proposal_get(E.iptr(cand), E.iptr(mode), None, P.N) == 0
            n_sh = L.emc_hip_proposal_shared(None, 0)
            shared = np.zeros(max(n_sh, 1), np.int32)
            assert L.emc_hip_proposal_shared(E.iptr(shared), n_sh) == n_sh
            sub = np.union1d(audited(P.N, it, frac), np.nonzero(mode_prev)[0])
            if st[14]:
                allowed = np.ones((P.R, P.N), bool)
            else:
                allowed = O.allowed_from_candidates(cand, off, rot, P.R, sub)
                for c in shared[:n_sh]:
                    allowed[rot[off[c]:off[c + 1]], :] = True
            ref = O.o64_proposal_iteration(state["model"].astype(np.float64), P.images, P.mask, state["scaling"],
                                           P.side, P.q, P.w, P.x, P.y, P.z, allowed, kind=O.POISSON)
            dP = np.abs(g["P"] - ref["P"]).max()
            ds = rel(g["scaling"], ref["scaling"])
            dm = rel(g["model"], ref["model"])
            dl = abs(res["likelihood"] - ref["total"]) / max(abs(ref["total"]), 1e-30)
            print(f"proposal iteration {it}: whole table {st[11]:.0f} (expected {len(sub)}), whole-table "
                  f"iteration {st[14]:.0f}, shared cells {n_sh}, full search {st[1]:.0f}, pairs {st[6]:.0f} "
                  f"(allowed {allowed.sum()}), audit fails {st[3]:.0f}; P {dP:.2e}, scaling {ds:.2e}, "
                  f"model {dm:.2e}, sum PlogP {dl:.2e}")
            if not st[14]:
                assert st[11] == len(sub) and st[6] == allowed.sum()
            assert dP < 1e-4 and ds < 1e-4 and dm < 1e-4
            worst = [max(worst[0], dP), max(worst[1], ds), max(worst[2], dm)]
            state = g
            mode_prev = mode
        print(f"worst over the proposal iterations: P {worst[0]:.2e}, scaling {worst[1]:.2e}, model {worst[2]:.2e}")
    finally:
        prob_.free()
        E.set_option("likelihood", "gaussian")
    E.check()
