"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Each entry point of emc.h is run on the GPU and compared with the oracle's
restatement of the reference on the same inputs:
  - expand (nearest voxel): bit-identical;
  - likelihood, scaling, M-step, insertion, normalisation: within the FP32
    rounding of the two evaluation orders (tolerances stated per test);
  - responsibilities: 1e-4 absolute; model: 1e-4 relative (the required
    agreement with the reference's method), one iteration at a time from
    the same inputs.
"""
import ctypes as C

import numpy as np
import pytest

import emc_amd as E
from oracle import oracle as O
from tests import problems

pytestmark = pytest.mark.gpu

P_TOL = 1e-4      # responsibilities, absolute
M_TOL = 1e-4      # model, relative


def options(precision="fp32", insert="atomic", interp="nearest"):
    E.set_option("precision", precision)
    E.set_option("insert", insert)
    E.set_option("interp", interp)


def rel_err(a, b, floor_frac=1e-3):
    """max relative error over entries above floor_frac * max|b|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sig = np.abs(b) > floor_frac * np.abs(b).max()
    return float(np.max(np.abs(a[sig] - b[sig]) / np.abs(b[sig]))) if sig.any() else 0.0


@pytest.fixture(scope="module")
def tiny():
    return problems.tiny()


@pytest.fixture(scope="module")
def cfg0():
    return problems.config0_like(n_patterns=1000)


@pytest.fixture(autouse=True)
def _reset():
    E.lib().emc_hip_clear_error()
    options()
    yield
    options()


def one_step(P, prob, sigma=0.4):
    """GPU steps one at a time, each compared with the FP32 restatement fed
    the GPU's own inputs of that step."""
    out = {}
    prob.get_slices()
    out["slices"] = prob.slices()
    Lh = np.zeros((prob.R, prob.N), np.float32)
    prob.likelihood(sigma, Lh)
    out["L"] = prob.respons()
    out["L_host"] = Lh
    prob.normalize_respons()
    out["P"] = prob.respons()
    out["total"] = prob.total_respons()
    prob.reset_model()
    prob.update_scaling()
    out["scaling"] = prob.scaling()
    out["scaling_host"] = prob.scaling_host.copy()
    out["overall"] = prob.update_slices()
    out["slices_new"] = prob.slices()
    out["model_raw"] = prob.model()
    out["weight"] = prob.weight()
    out["model_host_raw"] = prob.model_host.copy()
    prob.normalize_model()
    out["model"] = prob.model()
    return out


def check_one_step(P, prob, out, sigma=0.4, l_rtol=2e-5, fp16=False):
    model_in = P.model
    S = P.side
    sl = O.r32_expand(model_in, S, P.q, P.x, P.y, P.z)
    assert np.array_equal(out["slices"], sl), "expand must be bit-identical"
    sc_in = np.ones(P.N, np.float32)
    L = O.r32_likelihood(sl, P.images, P.mask, sigma, sc_in)
    assert np.array_equal(out["L"], out["L_host"])
    scale = np.abs(L).max()
    assert np.abs(out["L"] - L).max() <= l_rtol * scale, np.abs(out["L"] - L).max() / scale
    # responsibilities: GPU softmax of GPU L vs oracle softmax of oracle L;
    # against the FP64 responsibilities of the same inputs (the significant
    # entries are recomputed from FP64 likelihoods after the softmax)
    Pref = O.r32_softmax(L)
    assert np.abs(out["P"] - Pref).max() <= P_TOL
    L64 = O.o64_likelihood(sl.astype(np.float64), P.images, P.mask, sigma, sc_in.astype(np.float64))
    assert np.abs(out["P"] - O.o64_softmax(L64)).max() <= 2e-6
    t = O.r32_total_respons(out["P"])
    assert abs(out["total"] - t) <= 1e-5 * abs(t) + 1e-3
    phi = O.r32_update_scaling(P.images, sl, P.mask, out["P"])
    assert np.allclose(out["scaling"], phi, rtol=2e-5)
    assert np.array_equal(out["scaling"], out["scaling_host"])
    m = np.zeros(S ** 3, np.float32)
    wg = np.zeros(S ** 3, np.float32)
    ov, sl2, T = O.r32_update_slices(P.images, sl, P.mask, out["P"], out["scaling"], S, m, wg, P.x, P.y, P.z,
                                     P.q, P.w)
    assert abs(out["overall"] - ov) <= 1e-5 * abs(ov)
    msk = P.mask != 0
    assert rel_err(out["slices_new"][:, msk], sl2[:, msk]) <= (2e-4 if fp16 else 2e-5)
    assert np.array_equal(out["slices_new"][:, ~msk], sl[:, ~msk])  # unmasked pixels untouched
    assert rel_err(out["model_raw"], m) <= (1e-4 if fp16 else 2e-5)
    assert np.array_equal(out["model_raw"], out["model_host_raw"])
    assert rel_err(out["weight"], wg) <= 1e-5
    mn = O.r32_normalize_model(out["model_raw"], out["weight"])
    assert rel_err(out["model"], mn) <= 1e-6
    # end to end against the oracle's own chain from the same start
    ref_m, _, ref_s, ref_P, _ = O.r32_iteration(model_in, P.images, P.mask, sc_in, S, P.x, P.y, P.z, P.q, P.w,
                                                sigma)
    assert np.abs(out["P"] - ref_P).max() <= P_TOL
    assert rel_err(out["model"], ref_m) <= M_TOL


def test_expand_bitexact(tiny):
    options()
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    prob.get_slices()
    assert np.array_equal(prob.slices(), O.r32_expand(tiny.model, tiny.side, tiny.q, tiny.x, tiny.y, tiny.z))
    prob.free()


def test_one_step_fp32_tiny(tiny):
    options("fp32")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    out = one_step(tiny, prob)
    check_one_step(tiny, prob, out)
    prob.free()


def test_one_step_fp32_ragged():
    # N not a multiple of 4 (scalar load/store paths), R not a multiple of 128
    P = problems.tiny(n_patterns=37, n_rot=2, seed=5)
    options("fp32")
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q[:301], P.w[:301], P.model)
    Pq = problems.Problem(P.side, P.images, P.mask, P.q[:301].copy(), P.w[:301].copy(), P.x, P.y, P.z, P.model,
                          P.truth, P.phi_true)
    out = one_step(Pq, prob)
    check_one_step(Pq, prob, out)
    prob.free()


def test_one_step_fp16_tiny(tiny):
    options("fp16")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    out = one_step(tiny, prob)
    info = (C.c_int * 4)()
    E.lib().emc_hip_pack_info(info, 4)
    assert info[2] == 1  # integer counts are exact in FP16
    check_one_step(tiny, prob, out, fp16=True)
    prob.free()


def test_one_step_fp16_inexact_counts(tiny):
    # non-integer patterns force the split of K into hi + lo halves
    imgs = tiny.images * np.float32(1.37)
    P = problems.Problem(tiny.side, imgs, tiny.mask, tiny.q, tiny.w, tiny.x, tiny.y, tiny.z, tiny.model,
                         tiny.truth, tiny.phi_true)
    options("fp16")
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    out = one_step(P, prob)
    info = (C.c_int * 4)()
    E.lib().emc_hip_pack_info(info, 4)
    assert info[2] == 0 and info[3] == 3
    check_one_step(P, prob, out, fp16=True)
    prob.free()


def test_scaling_fallback_matches_fused(tiny):
    """update_scaling recomputes sum_p K S when the likelihood's partial sums
    are not available (the reference's method); both paths agree."""
    options("fp32")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    prob.get_slices()
    prob.likelihood(0.4)
    prob.normalize_respons()
    P = prob.respons()
    E.lib().emc_hip_invalidate()
    prob.update_scaling()
    phi_fb = prob.scaling()
    sl = prob.slices()
    ref = O.r32_update_scaling(tiny.images, sl, tiny.mask, P)
    assert np.allclose(phi_fb, ref, rtol=2e-5)
    prob.free()


def test_deterministic_insert_bit_reproducible(tiny):
    models = []
    for _ in range(2):
        options("fp32", "deterministic")
        prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
        for _ in range(2):
            prob.iteration()
        models.append((prob.model(), prob.weight()))
        prob.free()
        E.lib().emc_hip_invalidate()
    assert np.array_equal(models[0][0], models[1][0])
    assert np.array_equal(models[0][1], models[1][1])
    options("fp32", "atomic")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    for _ in range(2):
        prob.iteration()
    assert rel_err(prob.model(), models[0][0]) <= 1e-4
    prob.free()


@pytest.mark.parametrize("mode", ["deterministic", "deterministic_brick", "sorted_scatter"])
def test_deterministic_insert_matches_sequential_oracle(tiny, mode):
    """The deterministic forms add each voxel's contributions in the oracle's
    sequential order (rotation, then x, then y), so with the GPU's own
    updated slices and totals the accumulators agree bit for bit."""
    options("fp32", mode)
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    prob.get_slices()
    prob.likelihood(0.4)
    prob.normalize_respons()
    prob.reset_model()
    prob.update_scaling()
    Pm = prob.respons()
    prob.update_slices()
    sl2 = prob.slices()
    m_gpu, w_gpu = prob.model(), prob.weight()
    # the GPU's T_r is the float rounding of an FP64 sum of P over patterns
    T = Pm.astype(np.float64).sum(axis=1).astype(np.float32)
    S = tiny.side
    m = np.zeros(S ** 3, np.float32)
    wg = np.zeros(S ** 3, np.float32)
    lib = O.lib()
    for r in range(tiny.R):
        if T[r] > 1e-10:
            lib.r32_insert_slice(O.f(m), O.f(wg), S, O.f(np.ascontiguousarray(sl2[r])), O.i(tiny.mask),
                                 float(np.float32(tiny.w[r]) * T[r]), O.f(np.ascontiguousarray(tiny.q[r])),
                                 O.f(tiny.x), O.f(tiny.y), O.f(tiny.z))
    agree = np.mean(m_gpu == m)
    assert agree > 0.999, agree  # T may differ by one ulp for a handful of rotations
    assert rel_err(m_gpu, m) <= 1e-6
    assert rel_err(w_gpu, wg) <= 1e-6
    prob.free()


def test_trilinear_one_step_vs_fp64(tiny):
    options("fp32", "atomic", "trilinear")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    prob.iteration()
    ref = O.o64_iteration(tiny.model.astype(np.float64), tiny.images, tiny.mask, np.ones(tiny.N), tiny.side,
                          tiny.q, tiny.w, tiny.x, tiny.y, tiny.z, 0.4, O.TRILINEAR)
    assert np.abs(prob.respons() - ref["P"]).max() <= P_TOL
    assert rel_err(prob.model(), ref["model"]) <= M_TOL
    prob.free()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_config0_one_step(cfg0, precision):
    """Configuration 0's shapes (32^3, 1,000 patterns, 3,240 rotations)."""
    options(precision)
    prob = E.EmcProblem(cfg0.side, cfg0.images, cfg0.mask, cfg0.q, cfg0.w, cfg0.model)
    prob.iteration()
    ref_m, _, ref_s, ref_P, _ = O.r32_iteration(cfg0.model, cfg0.images, cfg0.mask, np.ones(cfg0.N, np.float32),
                                                cfg0.side, cfg0.x, cfg0.y, cfg0.z, cfg0.q, cfg0.w, 0.4)
    dp = np.abs(prob.respons() - ref_P).max()
    dm = rel_err(prob.model(), ref_m)
    print(f"config0 {precision}: max|dP| = {dp:.3g}, model rel = {dm:.3g}")
    assert dp <= P_TOL
    assert dm <= M_TOL
    assert np.allclose(prob.scaling(), ref_s, rtol=1e-4)
    prob.free()


def test_multi_iteration_tracks_oracle(tiny):
    """Three iterations; each GPU iteration is compared with the oracle's
    iteration started from the GPU's model and scaling of the step before."""
    options("fp32")
    prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
    model = tiny.model.copy()
    sc = np.ones(tiny.N, np.float32)
    for it in range(3):
        prob.iteration()
        ref_m, _, ref_s, ref_P, _ = O.r32_iteration(model, tiny.images, tiny.mask, sc, tiny.side, tiny.x, tiny.y,
                                                    tiny.z, tiny.q, tiny.w, 0.4)
        assert np.abs(prob.respons() - ref_P).max() <= P_TOL, it
        assert rel_err(prob.model(), ref_m) <= M_TOL, it
        model = prob.model()
        sc = prob.scaling()
    prob.free()


def test_error_reporting_invalid_arguments():
    L = E.lib()
    L.emc_hip_clear_error()
    L.cuda_normalize_responsabilities(None, 10, 10)
    buf = C.create_string_buffer(256)
    assert L.emc_hip_last_error(buf, 256) != 0
    assert b"invalid argument" in buf.value
    L.emc_hip_clear_error()
    assert L.emc_hip_last_error(buf, 256) == 0


@pytest.mark.parametrize("likelihood,interp", [("gaussian", "nearest"), ("poisson", "trilinear")])
def test_expand_pack_fused_matches_separate(tiny, likelihood, interp):
    """FP16 form: once an E-step has built the compact pixel list, the expand
    also packs the slice operands; the likelihood from those operands is
    bit-identical to the one from a separate pack of the same slices."""
    options("fp16", "atomic", interp)
    E.set_option("likelihood", likelihood)
    try:
        prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
        prob.iteration()  # builds the pixel list
        prob.get_slices()  # expand + pack
        sl_fused = prob.slices()
        L_fused = np.zeros((prob.R, prob.N), np.float32)
        prob.likelihood(0.4, L_fused)
        E.lib().emc_hip_invalidate()  # pixel list rebuilt: the E-step packs again
        prob.get_slices()
        sl_sep = prob.slices()
        L_sep = np.zeros((prob.R, prob.N), np.float32)
        prob.likelihood(0.4, L_sep)
        prob.free()
        E.check()
    finally:
        E.set_option("likelihood", "gaussian")
    assert np.array_equal(sl_fused, sl_sep)
    assert np.array_equal(L_fused, L_sep)


@pytest.mark.parametrize("side,n_patterns,n_rot", [(32, 600, 8), (16, 300, 9)])
def test_gemm_tail_split_matches_whole(side, n_patterns, n_rot):
    """FP16 likelihood GEMM with more tiles than compute units and a last
    round of at most half of them (600 x 25,680: 303 tiles; 300 x 36,540:
    286 tiles, 8 k steps, one per range): the tail tiles split along the
    pixel sum give the likelihood of the unsplit GEMM to FP32 rounding, and
    not bit for bit (the split ran)."""
    P = problems.make(side=side, n_patterns=n_patterns, n_rot=n_rot, n_blobs=20, radius=side / 5.0,
                      photons=1000.0, beamstop=2.0, seed=5)
    options("fp16")
    out = {}
    try:
        for mode in ("whole", "split"):
            E.set_option("gemm_tail", mode)
            prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
            prob.get_slices()
            L = np.zeros((prob.R, prob.N), np.float32)
            prob.likelihood(0.4, L)
            prob.free()
            E.check()
            out[mode] = L
    finally:
        E.set_option("gemm_tail", "split")
    w, s = out["whole"].astype(np.float64), out["split"].astype(np.float64)
    d = float(np.abs(s - w).max() / np.abs(w).max())
    print(f"{n_patterns} x {P.R}: max |L_split - L_whole| / max |L| = {d:.2e}, "
          f"{int((s != w).sum())} of {w.size} differ")
    assert d < 1e-6
    assert (s != w).any()


@pytest.mark.parametrize("likelihood,interp", [("gaussian", "nearest"), ("poisson", "trilinear")])
def test_expand_pack_large_slices_match_separate(likelihood, interp):
    """Slices too large for the LDS form (136 x 136): the expand writes the
    FP16 pair with each rotation's previous row scale and repacks the rows
    whose scale changed.  Over iterations from a random model (row scales
    changing) the slices are bit-identical to the separate pack's and the
    likelihood agrees to the rounding of A_r's summation order."""
    P = problems.make(side=136, n_patterns=64, n_rot=2, n_blobs=20, radius=24.0, photons=3000.0, beamstop=4.0,
                      seed=7)
    options("fp16", "atomic", interp)
    E.set_option("likelihood", likelihood)
    try:
        prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
        for it in range(3):
            prob.iteration()
            prob.get_slices()  # expand + pack with the previous row scales
            sl_f = prob.slices()
            L_f = np.zeros((prob.R, prob.N), np.float32)
            prob.likelihood(0.4, L_f)
            E.lib().emc_hip_invalidate()  # pixel list rebuilt: the E-step packs again
            prob.get_slices()
            sl_s = prob.slices()
            L_s = np.zeros((prob.R, prob.N), np.float32)
            prob.likelihood(0.4, L_s)
            E.check()
            d = float(np.abs(L_f.astype(np.float64) - L_s).max() / np.abs(L_s).max())
            print(f"iteration {it + 1}: slices equal {np.array_equal(sl_f, sl_s)}, max |dL| / max |L| = {d:.1e}")
            assert np.array_equal(sl_f, sl_s)
            assert d < 1e-6
        prob.free()
    finally:
        E.set_option("likelihood", "gaussian")


def test_gemm_tile_128_matches_256():
    """The 128 x 128 tile form of the FP16 likelihood GEMM (two workgroups per
    compute unit), the 256 x 128 form (four LDS stages) and the 8-wave
    256 x 256 form sum every output over k in the same order as the 16-wave
    256 x 256 form: bit-identical likelihoods (all unsplit along k)."""
    P = problems.make(side=32, n_patterns=300, n_rot=4, n_blobs=20, radius=6.0, photons=2000.0, seed=9)
    options("fp16")
    out = {}
    try:
        E.set_option("gemm_tail", "whole")
        for tile in ("256", "256w8", "128", "256x128"):
            E.set_option("gemm_tile", tile)
            prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
            prob.get_slices()
            L = np.zeros((prob.R, prob.N), np.float32)
            prob.likelihood(0.4, L)
            prob.free()
            E.check()
            out[tile] = L
    finally:
        E.set_option("gemm_tile", "256")
        E.set_option("gemm_tail", "split")
    assert np.array_equal(out["256"], out["128"])
    assert np.array_equal(out["256"], out["256x128"])
    assert np.array_equal(out["256"], out["256w8"])


@pytest.mark.parametrize("mask_kind", ["gap_h", "gap_cross"])
@pytest.mark.parametrize("precision,insert", [("fp32", "atomic"), ("fp32", "deterministic"), ("fp16", "atomic")])
def test_gap_masks_three_iterations(mask_kind, precision, insert):
    """The detector masks of the 64^3 workload (beamstop + 2-px horizontal
    gap) and of the 128^3 workloads (beamstop + 3-px cross), at 32^3 with
    3,240 rotations: three iterations, each against the FP64 oracle's and the
    FP32 restatement's iteration from the GPU's model and scaling of the step
    before; the model maximum the driver logs (cuda_model_max,
    emc_cuda.cu:364-368) against the oracle's max(0, max M) of the model it
    was computed from.  The FP32 restatement (the reference's own method) is
    itself off from the FP64 answer by up to ~2e-4 in P on these inputs (its
    FP32 sums of squares over ~800 pixels): the GPU must be within 1e-4 of
    the FP64 answer, and within 1e-4 of the restatement beyond the
    restatement's own error."""
    kw = dict(beamstop=4.0, gap_h=2) if mask_kind == "gap_h" else dict(beamstop=6.0, gap_cross=3)
    P = problems.make(side=32, n_patterns=300, n_rot=4, n_blobs=20, radius=6.0, photons=3000.0, phi_sigma=0.3,
                      seed=17, **kw)
    options(precision, insert)
    prob = E.EmcProblem(P.side, P.images, P.mask, P.q, P.w, P.model)
    model = P.model.copy()
    sc = np.ones(P.N, np.float32)
    for it in range(3):
        res = prob.iteration()
        assert res["model_max"] == O.r32_model_max(model), it
        ref_m, _, ref_s, ref_P, _ = O.r32_iteration(model, P.images, P.mask, sc, P.side, P.x, P.y, P.z, P.q, P.w,
                                                    0.4)
        o64 = O.o64_iteration(model.astype(np.float64), P.images, P.mask, sc.astype(np.float64), P.side, P.q, P.w,
                              P.x, P.y, P.z)
        g_P, g_m = prob.respons(), prob.model()
        dp, dm = np.abs(g_P - ref_P).max(), rel_err(g_m, ref_m)
        ep, em = np.abs(g_P - o64["P"]).max(), rel_err(g_m, o64["model"])
        rp, rm = np.abs(ref_P - o64["P"]).max(), rel_err(ref_m, o64["model"])
        print(f"{mask_kind} {precision} {insert} iteration {it + 1}: n_m {int((P.mask != 0).sum())}, "
              f"model max {res['model_max']:.4f}; vs FP64 P {ep:.2e} model {em:.2e}; vs FP32 restatement "
              f"P {dp:.2e} model {dm:.2e}; restatement vs FP64 P {rp:.2e} model {rm:.2e}")
        assert ep <= P_TOL and em <= M_TOL, it
        assert dp <= P_TOL + rp and dm <= M_TOL + rm, it
        assert np.allclose(prob.scaling(), o64["scaling"], rtol=1e-4), it
        model = prob.model()
        sc = prob.scaling()
    prob.free()


def test_expand_schedules_give_the_same_bits(tiny):
    """The expand's schedule options change which workgroup computes what,
    never a value: the fused LDS form with pixels in rows or 8 x 8 tiles and
    rotations in table or slice-normal order give the same slices and
    likelihoods, as do the rotation-block tiled forms with 4, 8 or 32
    rotations per block among themselves; between the two forms the slices
    are the same and the likelihoods differ by the order A_r is summed in.
    (Deterministic insertion, so that the model after the first iteration is
    the same bits in every variant.)"""
    options("fp16", "deterministic")
    out = {}
    variants = [("auto", "8", "rows", "normal"), ("auto", "8", "tiles", "normal"), ("auto", "8", "rows", "table"),
                ("tiled", "4", "rows", "normal"), ("tiled", "8", "rows", "table"), ("tiled", "32", "rows", "normal")]
    try:
        for v in variants:
            for k, val in zip(("expand_pack", "expand_rotations", "expand_pixels", "expand_order"), v):
                E.set_option(k, val)
            prob = E.EmcProblem(tiny.side, tiny.images, tiny.mask, tiny.q, tiny.w, tiny.model)
            prob.iteration()  # builds the pixel list, so the next expand packs
            prob.get_slices()
            L = np.zeros((prob.R, prob.N), np.float32)
            prob.likelihood(0.4, L)
            out[v] = (prob.slices(), L)
            prob.free()
            E.check()
            E.lib().emc_hip_invalidate()
    finally:
        for k, val in (("expand_pack", "auto"), ("expand_rotations", "8"), ("expand_pixels", "rows"),
                       ("expand_order", "normal")):
            E.set_option(k, val)
        options("fp32", "atomic")
    s0, L0 = out[variants[0]]
    t0 = out[variants[3]][1]
    for v in variants[1:]:
        assert np.array_equal(out[v][0], s0), v
        assert np.array_equal(out[v][1], L0 if v[0] == "auto" else t0), v
    d = float(np.abs(t0.astype(np.float64) - L0).max() / np.abs(L0).max())
    assert d < 1e-6, d
