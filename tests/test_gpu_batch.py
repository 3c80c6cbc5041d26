"""Ragged batches on the GPU: mixed sizes, modulations and operations in one
graph launch, checked transform by transform against the oracle; graph
replays and the per-size-launch baselines must reproduce the fused launch."""
import os

import numpy as np
import pytest

try:  # import at collection: the first import on a fresh box is slow
    import torch
except ImportError:  # pragma: no cover
    torch = None

import fftset_amd
from oracle import fp64
from oracle.oracle import MOD_COMPLEX as OCX, MOD_FREQ_OFFSET_REAL as OFO
from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

FO, CX = fftset_amd.FREQ_OFFSET_REAL, fftset_amd.COMPLEX
REF_TOL = 2e-6


@pytest.fixture(scope="module")
def fs():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    s = fftset_amd.FFTSet()
    yield s
    s.destroy()


def _make_jobs(rng, spec):
    """spec: list of (mod, L, op, count). Returns job list with host data."""
    jobs = []
    for mod, L, op, count in spec:
        for _ in range(count):
            x = rng.standard_normal(2 * L).astype(np.float32)
            h = (rng.standard_normal(2 * L) * np.exp(-np.arange(2 * L) / max(1.0, L / 4))).astype(np.float32)
            jobs.append(dict(mod=mod, L=L, op=op, x=x, h=h))
    order = rng.permutation(len(jobs))
    return [jobs[i] for i in order]


def _build(fs, jobs, flags=0):
    """Allocates device buffers, precomputes kernel spectra, adds to a batch."""
    b = fs.batch()
    for jb in jobs:
        f = fs.create_fft(jb["mod"], jb["L"])
        jb["f"] = f
        jb["dx"] = torch.from_numpy(jb["x"]).cuda()
        jb["dy"] = torch.full_like(jb["dx"], float("nan"))
        jb["dk"] = None
        if jb["op"] in (fftset_amd.OP_CONV, fftset_amd.OP_CONV_GRAD_INPUT):
            jb["dk"] = torch.empty_like(jb["dx"])
            f.conv_get_kernel(jb["dk"], torch.from_numpy(jb["h"]).cuda())
        b.add(f, jb["op"], jb["dy"], jb["dx"], jb["dk"])
    b.commit(flags)
    return b


def _oracle_ref(oracle, jb):
    o = oracle.plan(OFO if jb["mod"] == FO else OCX, jb["L"])
    op, x = jb["op"], jb["x"]
    if op == fftset_amd.OP_FORWARD:
        return o.forward(x)
    if op == fftset_amd.OP_INVERSE:
        return o.inverse(x)
    if op == fftset_amd.OP_CONV:
        return o.conv(x, o.get_kernel(jb["h"]))
    if op == fftset_amd.OP_CONV_GRAD_INPUT:
        # adjoint w.r.t. the signal: L * negacyclic (FREQOFFSETREAL) or L * cyclic (COMPLEX)
        # correlation with h, checked in FP64
        L = jb["L"]
        if jb["mod"] == FO:
            return fp64.fo_conv_grads64(x[None], np.zeros_like(x)[None], jb["h"][None])[0][0]
        X = np.fft.fft(fp64.as_complex(x))
        H = np.fft.fft(fp64.as_complex(jb["h"]))
        return fp64.as_interleaved(np.fft.ifft(X * np.conj(H)) * L)
    if op == fftset_amd.OP_GET_KERNEL:
        return None
    return None


SPEC = [
    (CX, 48, fftset_amd.OP_FORWARD, 40), (FO, 48, fftset_amd.OP_CONV, 40),
    (CX, 120, fftset_amd.OP_FORWARD, 20), (FO, 120, fftset_amd.OP_CONV, 20),
    (FO, 360, fftset_amd.OP_FORWARD, 9), (CX, 360, fftset_amd.OP_INVERSE, 9),
    (CX, 768, fftset_amd.OP_FORWARD, 7), (FO, 768, fftset_amd.OP_CONV, 7),
    (CX, 1440, fftset_amd.OP_CONV, 5), (FO, 1440, fftset_amd.OP_INVERSE, 5),
    (CX, 4608, fftset_amd.OP_FORWARD, 4), (FO, 4608, fftset_amd.OP_CONV, 4),
    (CX, 15360, fftset_amd.OP_FORWARD, 3), (FO, 15360, fftset_amd.OP_CONV, 3),
    (CX, 61440, fftset_amd.OP_FORWARD, 2), (FO, 61440, fftset_amd.OP_CONV, 2),
    (FO, 61440, fftset_amd.OP_INVERSE, 1), (CX, 61440, fftset_amd.OP_CONV, 1),
    (FO, 768, fftset_amd.OP_CONV_GRAD_INPUT, 3), (CX, 360, fftset_amd.OP_CONV_GRAD_INPUT, 3),
    (FO, 61440, fftset_amd.OP_CONV_GRAD_INPUT, 1), (CX, 120, fftset_amd.OP_GET_KERNEL, 2),
    (FO, 61440, fftset_amd.OP_GET_KERNEL, 1),
]


@pytest.mark.parametrize("flags", [0, fftset_amd.BATCH_NO_GRAPH, fftset_amd.BATCH_THREADS_512,
                                   fftset_amd.BATCH_SCHED_SERIAL, fftset_amd.BATCH_SCHED_PHASED,
                                   fftset_amd.BATCH_PERSISTENT,
                                   fftset_amd.BATCH_PERSISTENT | fftset_amd.BATCH_NO_GRAPH,
                                   fftset_amd.BATCH_FOUR_STEP_TILES,
                                   fftset_amd.BATCH_FOUR_STEP_TILES | fftset_amd.BATCH_PERSISTENT])
def test_ragged_batch_matches_oracle(fs, oracle, flags):
    rng = np.random.default_rng(11 + flags)
    jobs = _make_jobs(rng, SPEC)
    b = _build(fs, jobs, flags)
    b.launch()
    torch.cuda.synchronize()
    info = b.info()
    assert info["classes"] == len({(m, L, op) for m, L, op, _ in SPEC})
    for jb in jobs:
        y = jb["dy"].cpu().numpy()
        assert np.all(np.isfinite(y)), (jb["mod"], jb["L"], jb["op"])
        if jb["op"] == fftset_amd.OP_GET_KERNEL:
            # opaque spectrum: equal to the drop-in call on the same plan
            k = torch.empty_like(jb["dx"])
            jb["f"].conv_get_kernel(k, jb["dx"])
            assert fp64.rel_rms(y, k.cpu().numpy()) <= 1e-6
            continue
        ref = _oracle_ref(oracle, jb)
        tol = REF_TOL if jb["op"] != fftset_amd.OP_CONV_GRAD_INPUT else 1e-6 * np.log2(jb["L"])
        assert fp64.rel_rms(y, ref) <= tol, (jb["mod"], jb["L"], jb["op"])
    b.destroy()


@pytest.mark.parametrize("flags", [0, fftset_amd.BATCH_FOUR_STEP_TILES])
def test_graph_replays_and_baselines_reproduce_the_launch(fs, flags):
    rng = np.random.default_rng(5)
    spec = [(CX, 48, fftset_amd.OP_FORWARD, 16), (FO, 768, fftset_amd.OP_CONV, 6),
            (FO, 4608, fftset_amd.OP_CONV, 2), (CX, 61440, fftset_amd.OP_FORWARD, 3),
            (FO, 61440, fftset_amd.OP_CONV, 2)]
    jobs = _make_jobs(rng, spec)
    b = _build(fs, jobs, flags)
    b.launch()
    torch.cuda.synchronize()
    first = [jb["dy"].cpu().numpy().copy() for jb in jobs]
    for rep in range(3):
        for jb in jobs:  # eager work between replays
            jb["dy"].fill_(float("nan"))
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        for jb, ref in zip(jobs, first):
            assert np.array_equal(jb["dy"].cpu().numpy(), ref), rep
    for mode in (fftset_amd.BASELINE_PER_GROUP, fftset_amd.BASELINE_PER_TRANSFORM):
        for jb in jobs:
            jb["dy"].fill_(float("nan"))
        b.launch_baseline(mode)
        torch.cuda.synchronize()
        for jb, ref in zip(jobs, first):
            assert np.array_equal(jb["dy"].cpu().numpy(), ref), mode
    b.destroy()


def test_per_pass_baseline_matches_the_fused_launch(fs, oracle):
    # the reference's pass structure (one launch per radix pass, HBM between
    # passes) computes the same transforms as the fused LDS kernels
    rng = np.random.default_rng(9)
    jobs = _make_jobs(rng, SPEC)
    b = _build(fs, jobs)
    b.launch()
    torch.cuda.synchronize()
    first = [jb["dy"].cpu().numpy().copy() for jb in jobs]
    for rep in range(2):
        for jb in jobs:
            jb["dy"].fill_(float("nan"))
        b.launch_baseline(fftset_amd.BASELINE_PER_PASS)
        torch.cuda.synchronize()
        for jb, fused in zip(jobs, first):
            y = jb["dy"].cpu().numpy()
            assert np.all(np.isfinite(y)), (jb["mod"], jb["L"], jb["op"])
            assert fp64.rel_rms(y, fused) <= 1e-6, (jb["mod"], jb["L"], jb["op"])
            if jb["op"] in (fftset_amd.OP_FORWARD, fftset_amd.OP_INVERSE, fftset_amd.OP_CONV):
                assert fp64.rel_rms(y, _oracle_ref(oracle, jb)) <= REF_TOL, (jb["mod"], jb["L"], jb["op"])
    b.destroy()


REF_SPEC = SPEC + [
    # widths the ragged spec leaves out: width-1 COMPLEX (L % 4 != 0, and
    # L = 4), width-1 FREQOFFSETREAL inverse, width-4 FREQOFFSETREAL
    (CX, 90, fftset_amd.OP_FORWARD, 3), (CX, 90, fftset_amd.OP_CONV, 3), (CX, 45, fftset_amd.OP_INVERSE, 2),
    (CX, 4, fftset_amd.OP_FORWARD, 3), (FO, 360, fftset_amd.OP_INVERSE, 2), (FO, 240, fftset_amd.OP_CONV, 3),
    (FO, 80, fftset_amd.OP_FORWARD, 3), (FO, 240, fftset_amd.OP_CONV_GRAD_INPUT, 2),
]


def test_reference_pass_baseline_computes_the_same_transforms(fs):
    # the reference's pass structure (first pass, its planner's chain at its
    # vector width, product sweep, last pass; one launch per sweep) gives the
    # fused launch's results to FP32 rounding and matches FP64; get_kernel
    # jobs keep their spectra in the baseline's own order (not compared)
    rng = np.random.default_rng(21)
    jobs = _make_jobs(rng, REF_SPEC)
    b = _build(fs, jobs)
    b.launch()
    torch.cuda.synchronize()
    fused = [jb["dy"].cpu().numpy().copy() for jb in jobs]
    for rep in range(2):
        for jb in jobs:
            jb["dy"].fill_(float("nan"))
        b.launch_baseline(fftset_amd.BASELINE_REFERENCE_PASSES)
        torch.cuda.synchronize()
        for jb, fy in zip(jobs, fused):
            if jb["op"] == fftset_amd.OP_GET_KERNEL:
                continue
            tag = (jb["mod"], jb["L"], jb["op"], rep)
            y = jb["dy"].cpu().numpy()
            assert np.all(np.isfinite(y)), tag
            assert fp64.rel_rms(y, fy) <= REF_TOL, tag
            x, h, L = jb["x"], jb["h"], jb["L"]
            if jb["op"] == fftset_amd.OP_FORWARD:
                ref = fp64.fo_forward64(x) if jb["mod"] == FO else fp64.cx_forward64(x)
            elif jb["op"] == fftset_amd.OP_CONV:
                ref = fp64.fo_conv64(x, h) if jb["mod"] == FO else fp64.cx_conv64(x, h)
            else:
                continue
            assert fp64.rel_rms(y, ref) <= 1e-6 * max(1.0, np.log2(L)), tag
    b.destroy()


GK_SPEC = [
    (FO, 768, fftset_amd.OP_CONV_GRAD_KERNEL, 5), (CX, 360, fftset_amd.OP_CONV_GRAD_KERNEL, 4),
    (FO, 2048, fftset_amd.OP_CONV_GRAD_KERNEL, 3), (FO, 61440, fftset_amd.OP_CONV_GRAD_KERNEL, 2),
    (FO, 15360, fftset_amd.OP_CONV_GRAD_KERNEL, 1), (CX, 61440, fftset_amd.OP_CONV_GRAD_KERNEL, 1),
    (FO, 768, fftset_amd.OP_CONV, 4), (FO, 768, fftset_amd.OP_GET_KERNEL, 2), (CX, 4608, fftset_amd.OP_FORWARD, 3),
    (FO, 61440, fftset_amd.OP_CONV, 1),
]


def _grad_kernel64(jb):
    # dLoss/dh for output gradient g = the job's h-array and signal x
    x, g, L = jb["x"], jb["h"], jb["L"]
    if jb["mod"] == FO:
        return fp64.fo_conv_grads64(g[None], x[None], np.zeros_like(x)[None])[1][0]
    X = np.fft.fft(fp64.as_complex(x))
    G = np.fft.fft(fp64.as_complex(g))
    return fp64.as_interleaved(np.fft.ifft(G * np.conj(X)) * L)


def test_kernel_gradient_jobs_in_ragged_batches(fs):
    # FFTSET_OP_CONV_GRAD_KERNEL in a batch (in = signal, kern = output
    # gradient): its get_kernel stage and its adjoint stage run in one
    # launch; outputs match fftset_fft_execute and FP64, over graph replays
    # and every baseline
    rng = np.random.default_rng(33)
    jobs = _make_jobs(rng, GK_SPEC)
    b = fs.batch()
    for jb in jobs:
        f = fs.create_fft(jb["mod"], jb["L"])
        jb["f"] = f
        jb["dx"] = torch.from_numpy(jb["x"]).cuda()
        jb["dy"] = torch.full_like(jb["dx"], float("nan"))
        jb["dk"] = None
        if jb["op"] == fftset_amd.OP_CONV_GRAD_KERNEL:
            jb["dk"] = torch.from_numpy(jb["h"]).cuda()  # the output gradient
        elif jb["op"] == fftset_amd.OP_CONV:
            jb["dk"] = torch.empty_like(jb["dx"])
            f.conv_get_kernel(jb["dk"], torch.from_numpy(jb["h"]).cuda())
        b.add(f, jb["op"], jb["dy"], jb["dx"], jb["dk"])
    b.commit()
    expect = []
    for jb in jobs:
        ref = torch.full_like(jb["dx"], float("nan"))
        jb["f"].execute(jb["op"], 1, ref, jb["dx"], kern=jb["dk"])
        torch.cuda.synchronize()
        expect.append(ref.cpu().numpy())
    gk_bytes = sum(24 * jb["L"] for jb in jobs if jb["op"] == fftset_amd.OP_CONV_GRAD_KERNEL)
    assert b.bytes == gk_bytes + sum((24 if jb["op"] == fftset_amd.OP_CONV else 16) * jb["L"]
                                     for jb in jobs if jb["op"] != fftset_amd.OP_CONV_GRAD_KERNEL)
    for mode in ("fused", "fused", fftset_amd.BASELINE_PER_GROUP, fftset_amd.BASELINE_PER_TRANSFORM,
                 fftset_amd.BASELINE_PER_PASS, fftset_amd.BASELINE_REFERENCE_PASSES):
        for jb in jobs:
            jb["dy"].fill_(float("nan"))
        if mode == "fused":
            b.launch()
        else:
            b.launch_baseline(mode)
        torch.cuda.synchronize()
        for jb, ref in zip(jobs, expect):
            if jb["op"] == fftset_amd.OP_GET_KERNEL and mode == fftset_amd.BASELINE_REFERENCE_PASSES:
                continue  # that baseline keeps spectra in its own order
            tag = (mode, jb["mod"], jb["L"], jb["op"])
            y = jb["dy"].cpu().numpy()
            assert np.all(np.isfinite(y)), tag
            assert fp64.rel_rms(y, ref) <= (2e-6 if mode == fftset_amd.BASELINE_REFERENCE_PASSES else 1e-6), tag
            if jb["op"] == fftset_amd.OP_CONV_GRAD_KERNEL:
                assert fp64.rel_rms(y, _grad_kernel64(jb)) <= 1e-6 * np.log2(jb["L"]), tag
    b.destroy()


def test_batch_four_step_classes_equal_drop_in_calls(fs):
    # four-step classes run the drop-in four-step kernels through the
    # class's pointer arrays: outputs equal single-transform drop-in calls bit
    # for bit, whatever the order and addresses of the transforms
    rng = np.random.default_rng(17)
    spec = [(CX, 61440, fftset_amd.OP_FORWARD, 2), (FO, 61440, fftset_amd.OP_CONV, 2),
            (FO, 61440, fftset_amd.OP_INVERSE, 1), (CX, 61440, fftset_amd.OP_CONV_GRAD_INPUT, 1),
            (FO, 61440, fftset_amd.OP_GET_KERNEL, 1), (CX, 768, fftset_amd.OP_FORWARD, 3)]
    jobs = _make_jobs(rng, spec)
    b = _build(fs, jobs)
    b.launch()
    torch.cuda.synchronize()
    for jb in jobs:
        if jb["L"] != 61440:
            continue
        ref = torch.full_like(jb["dx"], float("nan"))
        jb["f"].execute(jb["op"], 1, ref, jb["dx"], kern=jb["dk"])
        torch.cuda.synchronize()
        assert np.array_equal(jb["dy"].cpu().numpy(), ref.cpu().numpy()), (jb["mod"], jb["op"])
    b.destroy()


def test_persistent_graph_replays_match_one_tile_per_workgroup(fs):
    # the tile counters are reset inside the graph: every replay redoes
    # every tile, and the results equal the one-tile-per-workgroup launch
    rng = np.random.default_rng(21)
    jobs = _make_jobs(rng, SPEC)
    b = _build(fs, jobs)
    b.launch()
    torch.cuda.synchronize()
    ref = [jb["dy"].cpu().numpy().copy() for jb in jobs]
    b.destroy()
    bp = fs.batch()
    for jb in jobs:
        bp.add(jb["f"], jb["op"], jb["dy"], jb["dx"], jb["dk"])
    bp.commit(fftset_amd.BATCH_PERSISTENT)
    for rep in range(3):
        for jb in jobs:
            jb["dy"].fill_(float("nan"))
        torch.cuda.synchronize()
        bp.launch()
        torch.cuda.synchronize()
        for jb, r in zip(jobs, ref):
            assert np.array_equal(jb["dy"].cpu().numpy(), r), (rep, jb["mod"], jb["L"], jb["op"])
    bp.destroy()


def test_batch_rejects_bad_use(fs):
    b = fs.batch()
    f = fs.create_fft(CX, 64)
    x = torch.zeros(128, device="cuda")
    with pytest.raises(fftset_amd.FFTSetError):
        b.add(f, fftset_amd.OP_CONV_GRAD_KERNEL, x, x, None)  # needs the output gradient
    with pytest.raises(fftset_amd.FFTSetError):
        b.add(f, fftset_amd.OP_CONV_SAVE_SPECTRUM, x, x, x)  # execute() only
    with pytest.raises(fftset_amd.FFTSetError):
        b.add(f, fftset_amd.OP_CONV, x, x, None)  # conv needs a kernel spectrum
    b.add(f, fftset_amd.OP_FORWARD, x, x)
    b.commit()
    with pytest.raises(fftset_amd.FFTSetError):
        b.commit()
    with pytest.raises(fftset_amd.FFTSetError):
        b.add(f, fftset_amd.OP_FORWARD, x, x)
    b.destroy()


def test_ragged_batch_takes_the_tuned_packing(oracle):
    # under the MI355X tune table every LDS class is packed as its plan is:
    # 32 transforms of 48 bins per 128 lanes, 42 of 120 bins per 512 lanes,
    # one of 1440 bins per 128 lanes, ...; the batch's largest tile is then
    # 42 x 120 bins (LDS rows of 120 + 120 / 16 + 1 complex values). Each
    # output equals the drop-in call on the same plan to rounding, and the
    # oracle.
    if torch_cuda() is None:
        pytest.skip("no GPU")
    tune = os.path.join(os.path.dirname(fftset_amd.__file__), "tune_mi355x.txt")
    s = fftset_amd.FFTSet(tune_file=tune)
    try:
        fwd, conv, inv = fftset_amd.OP_FORWARD, fftset_amd.OP_CONV, fftset_amd.OP_INVERSE
        spec = [(CX, 48, fwd, 40), (FO, 48, conv, 40), (CX, 120, fwd, 50), (FO, 120, conv, 45),
                (CX, 360, fwd, 12), (FO, 360, conv, 11), (CX, 768, fwd, 5), (FO, 768, conv, 6),
                (CX, 1440, fwd, 3), (FO, 1440, conv, 3), (CX, 1440, inv, 2)]
        jobs = _make_jobs(np.random.default_rng(23), spec)
        b = _build(s, jobs)
        assert b.info()["lds_bytes"] == 42 * (120 + 120 // 16 + 1) * 8
        b.launch()
        torch.cuda.synchronize()
        for jb in jobs:
            y = jb["dy"].cpu().numpy()
            drop_in = torch.full_like(jb["dx"], float("nan"))
            jb["f"].execute(jb["op"], 1, drop_in, jb["dx"], kern=jb["dk"])
            torch.cuda.synchronize()
            assert fp64.rel_rms(y, drop_in.cpu().numpy()) <= 2e-6, (jb["mod"], jb["L"], jb["op"])
            assert fp64.rel_rms(y, _oracle_ref(oracle, jb)) <= REF_TOL, (jb["mod"], jb["L"], jb["op"])
        b.destroy()
    finally:
        s.destroy()
