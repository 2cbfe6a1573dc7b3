"""Batched renders (pt_render_batch): many jobs -- different scenes, views,
image sizes, spp, seeds -- in one call over one merged device scene. Every
job's image, camera rays walked per ray in both (PT_RENDER_PER_RAY), must be
bit for bit the image pt_render gives for that job alone (which the oracle and
the reference fixtures pin), in shared-queue passes
(slices of many jobs in one bounce loop) or per-job loops over streams,
whatever the pass or batch sizes, streams and launch shape, with or without
the captured graphs, and after the autotuner."""
import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc

pytestmark = pytest.mark.gpu


def batch_scenes():
    return [rc.load_scene("field")[1], pt.pack(pt.generate_job(3, 0.01).ptsd), pt.pack(pt.generate_job(11, 0.01).ptsd),
            pt.pack(pt.generate(4, 2, 0.004)), rc.load_scene("materials")[1]]


def other_view(ptgp):
    """Camera rows of the scene's camera moved sideways and tilted: a second view of the same scene."""
    cam = np.frombuffer(ptgp[22:38].tobytes(), np.float32).reshape(4, 4).copy()
    cam[0, :3] += 0.35 * cam[2, :3]
    cam[1, :3] += 0.35 * cam[2, :3] + 0.1 * cam[3, :3]
    return cam


def with_camera(ptgp, cam):
    w = ptgp.copy()
    w[22:38] = np.frombuffer(np.asarray(cam, np.float32).tobytes(), np.uint32)
    return w


def jobs_of(scenes):
    return [
        pt.BatchJob(0, 96, 54, 16, seed=5),
        pt.BatchJob(1, 64, 64, 4, seed=6),
        pt.BatchJob(2, 80, 45, 64, seed=7, first_sample=3),
        pt.BatchJob(3, 72, 40, 16, seed=8),
        pt.BatchJob(0, 48, 48, 8, seed=9, nee=False),
        pt.BatchJob(4, 90, 50, 4, seed=10, camera=other_view(scenes[4])),
        pt.BatchJob(1, 33, 17, 12, seed=11),
    ]


def single(scenes, j, depth=10):
    ptgp = scenes[j.scene] if j.camera is None else with_camera(scenes[j.scene], j.camera)
    with pt.Scene(ptgp) as s:
        img, _, st = s.render(j.width, j.height, j.spp, max_bounces=depth, seed=j.seed, first_sample=j.first_sample,
                              nee=j.nee, per_ray=True)
    return img, st


@pytest.fixture(scope="module")
def singles():
    sc = batch_scenes()
    jobs = jobs_of(sc)
    return sc, jobs, [single(sc, j) for j in jobs]


@pytest.mark.parametrize("shared", [True, False])
def test_batch_equals_single_renders(singles, shared):
    """Shared-queue passes (the default) and per-job bounce loops over streams."""
    sc, jobs, ref = singles
    with pt.Batch(sc) as b:
        imgs, st = b.render(jobs, shared=shared, per_ray=True)
    for k, (img, (want, _)) in enumerate(zip(imgs, ref)):
        assert img.shape == want.shape
        same = np.array_equal(img.view(np.uint32), want.view(np.uint32))
        print(f"job {k}: {'bit-identical' if same else 'DIFFERS'}; mean {img.mean():.5f}", flush=True)
        assert same, k
    assert st.rays == sum(s.rays for _, s in ref)
    assert st.primary_rays == sum(s.primary_rays for _, s in ref)
    assert st.shadow_rays == sum(s.shadow_rays for _, s in ref)


@pytest.mark.parametrize("max_paths,streams,shared", [(6000, 1, False), (40000, 3, False), (1 << 20, 2, False),
                                                      (1 << 20, 8, False), (6000, 1, True), (9000, 1, True),
                                                      (40000, 1, True), (1 << 20, 1, True)])
def test_batch_sizes_and_streams_change_nothing(singles, max_paths, streams, shared):
    """Small queues split every job into many batches (per-job loops) or passes whose slices split jobs
    between them (shared queues); jobs spread over several streams; eager launches and the graphs agree."""
    sc, jobs, ref = singles
    with pt.Batch(sc) as b:
        a, sa = b.render(jobs, max_paths=max_paths, streams=streams, shared=shared, per_ray=True)
        e, se = b.render(jobs, max_paths=max_paths, streams=streams, graph=False, shared=shared, per_ray=True)
    assert sa.batches == se.batches
    if max_paths == 6000:
        assert sa.batches > 3 * len(jobs) if not shared else sa.batches > 40
    if shared and max_paths == 1 << 20:
        assert sa.batches == 1
    for x, y, (want, _) in zip(a, e, ref):
        assert np.array_equal(x, want) and np.array_equal(y, want)
    assert sa.rays == se.rays == sum(s.rays for _, s in ref)


@pytest.mark.parametrize("shared", [True, False])
def test_device_output_and_stage_by_stage(singles, shared):
    sc, jobs, ref = singles
    n = sum(j.width * j.height for j in jobs)
    buf = pt.DeviceBuffer(12 * n)
    with pt.Batch(sc) as b:
        _, sd = b.render(jobs, device_out=buf, shared=shared, per_ray=True)
        flat = buf.download(np.zeros(3 * n, np.float32))
        nt, _ = b.render(jobs, tail=False, shared=shared, per_ray=True)
    at = 0
    for j, x, (want, _) in zip(jobs, nt, ref):
        k = j.width * j.height
        assert np.array_equal(flat[3 * at:3 * (at + k)].reshape(want.shape), want)
        assert np.array_equal(x, want)
        at += k
    assert sd.rays == sum(s.rays for _, s in ref)


def test_launch_shapes_change_nothing(singles):
    sc, jobs, ref = singles
    with pt.Batch(sc) as b:
        for S, per_cu, shared in ((16, 0, True), (32, 2, True), (8, 3, True), (16, 0, False), (8, 3, False)):
            imgs, _ = b.render(jobs, stack_entries=S, blocks_per_cu=per_cu, shared=shared, per_ray=True)
            for x, (want, _) in zip(imgs, ref):
                assert np.array_equal(x, want), (S, per_cu, shared)
        # the two modes one after the other on the same batch (each keeps its own captured graph)
        for shared in (True, False, True):
            imgs, _ = b.render(jobs, shared=shared, per_ray=True)
            for x, (want, _) in zip(imgs, ref):
                assert np.array_equal(x, want), shared


@pytest.mark.parametrize("shared", [True, False])
def test_batch_camera_packets_agree_with_per_ray(singles, shared):
    """Batched camera rays walked as 64-ray packets (the default; a run of the shared queue holding paths of
    two jobs is walked once per job) against the same batch walked per ray: each job's image bit for bit
    equal on at least 99.99% of its pixels (the tie bound), the ray counts to 1e-4."""
    sc, jobs, _ = singles
    with pt.Batch(sc) as b:
        pk, sp = b.render(jobs, shared=shared)
        pr, sr = b.render(jobs, shared=shared, per_ray=True)
    for k, (x, y) in enumerate(zip(pk, pr)):
        same = np.all(x.view(np.uint32) == y.view(np.uint32), axis=2)
        print(f"job {k}: {int((~same).sum())} of {same.size} pixels differ", flush=True)
        assert same.mean() >= 0.9999, k
    assert abs(sp.rays - sr.rays) <= 1e-4 * sr.rays and sp.primary_rays == sr.primary_rays


def test_batch_job_matches_oracle():
    """A batched job against the CPU oracle directly (same tolerance as the single-render test)."""
    sc = batch_scenes()
    job = pt.BatchJob(3, 96, 54, 8, seed=11)
    with pt.Batch(sc) as b:
        (_, img), st = b.render([pt.BatchJob(0, 40, 30, 4, seed=1), job])
    orgb, _, _ = ob.or_render(sc[3], 96, 54, 8, seed=11)
    frac, note = rc.per_value_agreement(img, orgb)
    print(f"{frac:.5f} of values within 1e-4 ({note})", flush=True)
    assert frac >= 0.999


def test_autotune_then_render(singles):
    sc, jobs, ref = singles
    with pt.Batch(sc) as b:
        r = b.autotune(jobs)
        print(f"tuned: stack {r.stack_entries}, blocks/CU {r.blocks_per_cu}, streams {r.streams}, paths {r.max_paths}; "
              f"tried {[(t[0], t[1], t[2], t[3], round(t[4])) for t in r.tried]}", flush=True)
        assert r.stack_entries in (8, 16, 32) and r.blocks_per_cu >= 1 and 1 <= r.streams <= 8
        assert len(r.tried) >= 6
        imgs, _ = b.render(jobs, per_ray=True)
    for x, (want, _) in zip(imgs, ref):
        assert np.array_equal(x, want)


def test_bad_batches_are_refused():
    sc = batch_scenes()
    with pt.Batch(sc[:2]) as b:
        for bad in (pt.BatchJob(2, 32, 32, 4), pt.BatchJob(0, 1, 32, 4), pt.BatchJob(0, 32, 32, 0),
                    pt.BatchJob(0, 32, 32, 4, first_sample=-1)):
            with pytest.raises(pt.PTError):
                b.render([bad])
        with pytest.raises(pt.PTError):
            b.render([pt.BatchJob(0, 32, 32, 4)], max_bounces=0)
        with pytest.raises(pt.PTError):
            b.render([pt.BatchJob(0, 64, 64, 4)], max_paths=1000)   # a job's pixels must fit one launch
        imgs, _ = b.render([pt.BatchJob(0, 32, 32, 4)])              # still usable
        assert imgs[0].shape == (32, 32, 3)
    with pytest.raises(pt.PTError):
        pt.Batch([sc[0], pt.pack(pt.generate(2, 0, 0.002))])
