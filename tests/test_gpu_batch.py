"""Batched uneven-size DTS (one graph replay for all frames) against the FP64
oracle frame by frame, and against the single-frame path."""
import numpy as np
import pytest

from domain_transform_solver_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


def _oracle(oracle_mod, f):
    h, w = f.target.shape
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(f.color_space, f.rgba, f.target, f.confidence)
    fp, op = f.filter_options, f.optim_options
    o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    o.integrate()
    o.optimize(op.num_iterations, op.loss, op.lambda_, op.sigma_z, op.step_size,
               f.overwrite_target_above_conf)
    return o.result()


def _depth_frame(S, w, h, factor, seed, K=10):
    # Frames narrower than one low-resolution block are cut from a wider scene.
    sc = scenes.depth_scene(max(w, factor), max(h, factor), factor, seed=seed)
    p = scenes.depth_sr_params(factor)
    return S.BatchFrameSpec(np.ascontiguousarray(sc.rgba[:h, :w]),
                            np.ascontiguousarray(sc.target[:h, :w]),
                            np.ascontiguousarray(sc.confidence[:h, :w]),
                            S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
                            S.DomainOptimizeParams(K, p["loss"], p["lam"], p["sigma_z"],
                                                   p["step"]), p["color_space"])


def test_batch_uneven_depth_frames(S, oracle_mod):
    specs = [(320, 240, 8), (200, 130, 4), (97, 301, 16), (640, 480, 8), (1, 50, 4), (129, 77, 4)]
    frames = [_depth_frame(S, w, h, f, 4000 + i) for i, (w, h, f) in enumerate(specs)]
    total = sum(w * h for w, h, _ in specs)
    b = S.BatchSolver(len(frames), total)
    b.SetFrames(frames)
    b.Solve()
    for i, f in enumerate(frames):
        ref = _oracle(oracle_mod, f)
        out = b.Download(i)
        rng = float(f.target.max() - f.target.min()) or 1.0
        err = float(np.abs(out - ref).max())
        print(f"frame {i} {f.target.shape}: max-abs {err:.3e} / range {rng:.3f}")
        assert err <= 1e-4 * rng


@pytest.mark.parametrize("nq", [8, 10])
def test_batch_distinct_sigma_ratios(S, oracle_mod, nq):
    """Frames with nq distinct range sigmas: up to 8 distinct ratios the
    bounds come from the guide and one step table per ratio, beyond that
    from the ct scan; both match the oracle frame by frame."""
    frames = []
    for i in range(nq):
        f = _depth_frame(S, 96 + 8 * i, 64 + 4 * (i % 3), 4, 4050 + i, K=4)
        fp = f.filter_options
        f.filter_options = S.DomainFilterParams(fp.sigma_x, fp.sigma_y, fp.sigma_r * (1 + 0.1 * i))
        frames.append(f)
    b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
    b.SetFrames(frames)
    b.Solve()
    for i, f in enumerate(frames):
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        err = float(np.abs(b.Download(i) - ref).max())
        print(f"frame {i} {f.target.shape}: max-abs {err:.3e} / range {rng:.3f}")
        assert err <= 1e-4 * rng


def test_batch_planned_by_measured_extents(S, oracle_mod):
    """Textured frames of factors 4, 8 and 16 have windows of a few pixels
    whatever their sigma, so all of them take the 2-chunk fused tiles; the
    same frames with flat guides need their full radii.  Both sets match
    the oracle."""
    tex = [_depth_frame(S, 192, 120, f, 6200 + i, K=4) for i, f in enumerate((4, 8, 16))]
    flat = [_flat_guide(_depth_frame(S, 192, 120, f, 6200 + i, K=4))
            for i, f in enumerate((4, 8, 16))]
    for frames, fused_at_2 in ((tex, 3), (flat, 1)):
        b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
        b.SetFrames(frames)
        b.SetFused(2)
        assert b.FramesFused() == fused_at_2
        for fz in (0, 2, 4):
            b.SetFused(fz)
            b.Solve()
            for i, f in enumerate(frames):
                ref = _oracle(oracle_mod, f)
                rng = float(f.target.max() - f.target.min()) or 1.0
                err = float(np.abs(b.Download(i) - ref).max())
                print(f"fused <= {fz} frame {i}: {err / rng:.2e} x range")
                assert err <= 1e-4 * rng
        b.close()


def test_batch_extent_planning_off_uses_caps(S, oracle_mod):
    """With extent planning off the passes are planned by the sigma-based
    caps (p - l <= floor(sigma) + 1, u - p <= floor(sigma)); on, by the
    measured extents, which are never larger.  Both solve to the oracle."""
    frames = [_depth_frame(S, 200, 150, f, 6300 + i, K=4) for i, f in enumerate((4, 16))]
    res = {}
    for on in (True, False):
        b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
        b.SetExtentPlanning(on)
        b.SetFrames(frames)
        ext = [b.FrameExtents(i) for i in range(len(frames))]
        b.Solve()
        res[on] = (ext, [b.Download(i) for i in range(len(frames))])
        b.close()
    for i, f in enumerate(frames):
        sz = f.optim_options.sigma_z
        cap = (int(sz) + 1, int(sz), int(sz) + 1, int(sz))
        assert res[False][0][i] == cap
        assert all(m <= c for m, c in zip(res[True][0][i], cap))
        assert max(res[True][0][i]) < 20 < cap[0] or i == 0
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        for on in (True, False):
            assert float(np.abs(res[on][1][i] - ref).max()) <= 1e-4 * rng


def test_batch_with_wide_bounds_frame(S, oracle_mod):
    """One frame with radius 280 (extents above 254) puts the whole batch on
    4-byte bounds elements; every frame still matches the oracle."""
    wide = _depth_frame(S, 320, 300, 8, 4100, K=4)
    wide.optim_options = S.DomainOptimizeParams(4, S.L2, 0.99, 280.0, 0.99)
    frames = [wide, _depth_frame(S, 200, 130, 4, 4101, K=4)]
    b = S.BatchSolver(len(frames), 320 * 300 + 200 * 130)
    b.SetFrames(frames)
    b.Solve()
    for i, f in enumerate(frames):
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        err = float(np.abs(b.Download(i) - ref).max())
        print(f"frame {i} {f.target.shape}: max-abs {err:.3e} / range {rng:.3f}")
        assert err <= 1e-4 * rng


def test_batch_frame_beyond_the_segments(S, oracle_mod):
    """Radius 300 on 1100x300 exceeds the batched kernels' segments: that frame
    is solved by the single-frame wide-window path in the same sequence, the
    others by the batched kernels; graph replays repeat the result."""
    big = _depth_frame(S, 1100, 300, 8, 4200, K=3)
    big.optim_options = S.DomainOptimizeParams(3, S.L2, 0.99, 300.0, 0.99)
    frames = [_depth_frame(S, 320, 240, 8, 4201, K=3), big, _depth_frame(S, 200, 130, 4, 4202, K=3)]
    b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
    b.SetFrames(frames)
    b.Solve()
    first = [b.Download(i) for i in range(len(frames))]
    b.Solve()
    for i, f in enumerate(frames):
        out = b.Download(i)
        np.testing.assert_array_equal(out, first[i])
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        err = float(np.abs(out - ref).max())
        print(f"frame {i} {f.target.shape}: max-abs {err:.3e} / range {rng:.3f}")
        assert err <= 1e-4 * rng


def test_batch_long_lines_use_workgroup_scan(S, oracle_mod):
    """A frame with rows longer than 2048 switches the batch's ct scan to one
    workgroup per line."""
    frames = [_depth_frame(S, 2100, 8, 4, 4300), _depth_frame(S, 64, 48, 4, 4301)]
    b = S.BatchSolver(2, 2100 * 8 + 64 * 48)
    b.SetFrames(frames)
    b.Solve()
    for i, f in enumerate(frames):
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        assert float(np.abs(b.Download(i) - ref).max()) <= 1e-4 * rng


def test_batch_mixed_options_and_graph_replays(S, oracle_mod):
    sc = scenes.colorization_scene(300, 200, seed=5, regions=6, scribble_frac=0.03)
    cp = scenes.colorization_params()
    col = S.BatchFrameSpec(sc.rgba, sc.target_u, sc.confidence,
                           S.DomainFilterParams(cp["sigma_x"], cp["sigma_y"], cp["sigma_r"]),
                           S.DomainOptimizeParams(7, cp["loss"], cp["lam"], cp["sigma_z"],
                                                  cp["step"]), 0, cp["overwrite"])
    d1 = _depth_frame(S, 256, 128, 8, 7, K=3)
    l1 = _depth_frame(S, 64, 64, 4, 8)
    l1.optim_options.loss = S.L1
    frames = [d1, col, l1]
    b = S.BatchSolver(4, 300 * 200 + 256 * 128 + 64 * 64)
    b.SetFrames(frames)
    first = None
    for rep in range(3):
        b.Solve()
        outs = [b.Download(i) for i in range(3)]
        if first is None:
            first = outs
        for o, f in zip(outs, first):
            np.testing.assert_array_equal(o, f)
        b.Synchronize()
    np.testing.assert_array_equal(first[2], l1.target)
    for i in (0, 1):
        ref = _oracle(oracle_mod, frames[i])
        rng = float(frames[i].target.max() - frames[i].target.min())
        assert float(np.abs(first[i] - ref).max()) <= 1e-4 * rng


def test_batch_matches_single_frame_path(S):
    f = _depth_frame(S, 320, 240, 8, 1000)
    b = S.BatchSolver(1, 320 * 240)
    b.SetFrames([f])
    b.Solve()
    gs = S.DomainTransformSolver(320, 240)
    gs.InitFrame(f.color_space, f.rgba, f.target, f.confidence)
    gs.ComputeColorSpaceDifferential(f.filter_options)
    gs.IntegrateColorDifferentials()
    gs.Optimize(f.optim_options)
    # Same segment and tile arithmetic, but the batched pass stages 8 chunks
    # per segment (9 elements per lane in the serial part of the LDS scan, 7
    # in the single-frame pass for this radius), so the fp32 prefix sums
    # associate differently; both paths meet the oracle bar on their own.
    rng = float(f.target.max() - f.target.min())
    np.testing.assert_allclose(b.Download(0), gs.Download(S.OPTIMIZED_QUANTITY), rtol=0,
                               atol=1e-5 * rng)


def _mixed_frames(S):
    specs = [(320, 240, 4), (200, 130, 8), (97, 301, 16), (640, 480, 8), (129, 77, 4)]
    return [_depth_frame(S, w, h, f, 5000 + i) for i, (w, h, f) in enumerate(specs)], \
        sum(w * h for w, h, _ in specs)


def test_batch_tunings_give_identical_results(S):
    """Every tuning choice of the two-pass iterations (persistent grid,
    register or LDS vertical tiles) performs the same arithmetic."""
    frames, total = _mixed_frames(S)
    b = S.BatchSolver(len(frames), total)
    b.SetFrames(frames)
    b.SetFused(0)
    ref = None
    for yr in (False, True):
        for xb in (1, 2, 4, 8):
            b.SetTuning(xb, yr)
            b.Solve()
            outs = [b.Download(i) for i in range(len(frames))]
            if ref is None:
                ref = outs
            for o, r in zip(outs, ref):
                np.testing.assert_array_equal(o, r)


def test_batch_autotune_and_cache(S, tmp_path):
    frames, total = _mixed_frames(S)
    cache = str(tmp_path / "tune.json")
    b = S.BatchSolver(len(frames), total)
    b.SetFrames(frames)
    t1 = b.Autotune(cache, reps=1, device="test")
    assert not t1["cached"] and t1["ms"] > 0
    b.Solve()
    out = [b.Download(i) for i in range(len(frames))]
    b2 = S.BatchSolver(len(frames), total)
    b2.SetFrames(frames)
    t2 = b2.Autotune(cache, reps=1, device="test")
    assert t2["cached"]
    assert b2.GetTuning() == b.GetTuning()
    b2.Solve()
    for i, o in enumerate(out):
        np.testing.assert_array_equal(b2.Download(i), o)


@pytest.mark.parametrize("waves", [4, 8])
def test_batch_fused_iterations(S, oracle_mod, waves):
    """One launch per iteration for the whole batch: frames in all three
    chunk classes (factors 4, 8, 16), odd and even iteration counts (each
    frame's last iterate must land in its result plane), a 1-wide frame;
    tiles of 4 or 8 wavefronts.  Every frame against the oracle and against
    the two-pass iterations; graph replays repeat the result bit for bit."""
    specs = [(320, 240, 4, 10), (200, 130, 8, 3), (97, 301, 16, 10), (640, 480, 8, 7),
             (1, 50, 4, 4), (129, 77, 16, 1), (300, 64, 4, 2)]
    frames = [_depth_frame(S, w, h, f, 6000 + i, K=k) for i, (w, h, f, k) in enumerate(specs)]
    b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
    b.SetFrames(frames)
    b.SetFused(0)
    assert b.FramesFused() == 0
    b.Solve()
    two_pass = [b.Download(i) for i in range(len(frames))]
    b.SetFused(4)
    b.SetFusedWaves(waves)
    assert b.FramesFused() == len(frames)
    assert b.GetTuning()["fused_waves"] == waves
    first = None
    for rep in range(3):
        b.Solve()
        outs = [b.Download(i) for i in range(len(frames))]
        if first is None:
            first = outs
        for o, f in zip(outs, first):
            np.testing.assert_array_equal(o, f)
    for i, f in enumerate(frames):
        ref = _oracle(oracle_mod, f)
        rng = float(f.target.max() - f.target.min()) or 1.0
        err = float(np.abs(first[i] - ref).max())
        d2 = float(np.abs(first[i] - two_pass[i]).max())
        print(f"frame {i} {f.target.shape} K={f.optim_options.num_iterations}: "
              f"vs oracle {err / rng:.2e}, vs two-pass {d2 / rng:.2e} x range")
        assert err <= 1e-4 * rng
        assert d2 <= 5e-5 * rng


def _flat_guide(f):
    """The frame with a uniform guide: every step is 1, so its windows reach
    the full sigma-based radius and its measured extents equal the caps."""
    f.rgba = np.full_like(f.rgba, 90)
    return f


def test_batch_fused_and_two_pass_frames_together(S, oracle_mod):
    """Frames with the target overwrite, and frames past the chunk limit,
    keep the two launches per iteration in the same replay as the fused
    frames.  The passes are planned by each frame's measured window
    extents: the textured factor-4 frame spans 2 chunks, the flat-guide
    factor-16 and factor-8 frames their full radii (4 and 3 chunks)."""
    sc = scenes.colorization_scene(200, 120, seed=9, regions=5, scribble_frac=0.03)
    cp = scenes.colorization_params()
    col = S.BatchFrameSpec(sc.rgba, sc.target_u, sc.confidence,
                           S.DomainFilterParams(cp["sigma_x"], cp["sigma_y"], cp["sigma_r"]),
                           S.DomainOptimizeParams(5, cp["loss"], cp["lam"], cp["sigma_z"],
                                                  cp["step"]), 0, cp["overwrite"])
    frames = [_depth_frame(S, 160, 120, 4, 6100, K=4), col,
              _flat_guide(_depth_frame(S, 200, 150, 16, 6101, K=3)),
              _flat_guide(_depth_frame(S, 150, 90, 8, 6102, K=5))]
    b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
    b.SetFrames(frames)
    refs = [_oracle(oracle_mod, f) for f in frames]
    for fz, nf in ((2, 1), (3, 2), (4, 3)):
        b.SetFused(fz)
        assert b.FramesFused() == nf
        b.Solve()
        for i, f in enumerate(frames):
            rng = float(f.target.max() - f.target.min()) or 1.0
            err = float(np.abs(b.Download(i) - refs[i]).max())
            print(f"chunks <= {fz}: frame {i} {f.target.shape}: {err / rng:.2e} x range")
            assert err <= 1e-4 * rng
