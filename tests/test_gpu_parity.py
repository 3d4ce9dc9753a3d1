"""HIP path vs the oracle on identical seeded inputs (needs an MI355X).

Gates: Sobol samples and BVH primitive order bit-exact; closest-hit primitive ids equal
on >= 99.99% of rays with mismatches only at near ties; radiance relative RMSE <= 1e-3
at equal sample indices; bakes relative RMSE <= 1e-3.
"""
import math

import numpy as np
import pytest

from axomae_amd import scenes
from conftest import rel_rmse
from scene_cases import BVH_SCENES, hit_test_rays, near_axis_test_rays, soup_scene

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from axomae_amd.nova_hip import HipEngine
    e = HipEngine(0)
    yield e
    e.close()


@gpu
def test_sobol_bit_exact(oracle):
    import torch
    from axomae_amd import nova_hip
    rng = np.random.default_rng(11)
    idx = np.concatenate([[0, 1, 2, 3, 3_932_100, 522_546_948, 0xFFFFFFFF], rng.integers(0, 2 ** 32, 4089, dtype=np.uint64)]).astype(np.uint32)
    I, D = np.meshgrid(idx, np.arange(256, dtype=np.uint32), indexing="ij")
    I, D = I.ravel(), D.ravel()
    out, bits = nova_hip.sobol_fill(torch.from_numpy(I.view(np.int32)), torch.from_numpy(D.view(np.int32)), return_bits=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint32), oracle.sobol_bits(I, D))
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.sobol_float(I, D))


@gpu
def test_sobol_equals_reference_generator_vectors():
    """nova_hip_sobol_fill against the integers and floats written by the reference's own
    scripts/generate_luts_sobol.py (tests/golden/sobol_ref_golden.npz), all 256 dims."""
    import torch
    from axomae_amd import nova_hip
    from test_sampler import reference_vectors
    g = reference_vectors()
    I, D = np.meshgrid(g["index"], np.arange(256, dtype=np.uint32), indexing="ij")
    out, bits = nova_hip.sobol_fill(torch.from_numpy(I.ravel().view(np.int32)), torch.from_numpy(D.ravel().view(np.int32)), return_bits=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint32).reshape(I.shape), g["bits"])
    keep = g["index"] != 0
    np.testing.assert_array_equal(out.cpu().numpy().reshape(I.shape)[keep][:, :255], g["value"][keep][:, :255])


def _leaf_set(nodes):
    return sorted((int(n["left"]), int(n["count"])) for n in nodes if n["count"] > 0)


@gpu
@pytest.mark.parametrize("name", sorted(BVH_SCENES))
def test_bvh_primitive_order_bit_exact(oracle, engine, name):
    sc = BVH_SCENES[name]()
    print(f"[bvh] {name}: {sc.triangle_count} triangles", flush=True)
    engine.upload(sc)
    info = engine.build_bvh()
    prim, nodes = engine.bvh_arrays()
    os_ = oracle.OracleScene(sc)
    oprim, onodes = os_.bvh_arrays()
    np.testing.assert_array_equal(prim, oprim)
    assert info.binary_nodes == len(onodes)
    assert _leaf_set(nodes) == _leaf_set(onodes)
    # inner-node boxes equal as a multiset (numbering differs: level order vs depth-first)
    key = lambda arr: sorted(tuple(np.concatenate([a["min"], a["max"]]).tolist()) + (int(a["count"]),) for a in arr)
    assert key(nodes) == key(onodes)


@gpu
@pytest.mark.parametrize("name", ["box_room", "terrain_spheres", "soup_5000", "soup_duplicates", "soup_int_max_fallback", "transformed"])
def test_closest_hit_matches_reference_traversal(oracle, engine, name):
    """Primitive ids, t, u and v identical to the reference walk on every ray (the gate is
    99.99% of ids; the traversal's tie rule and parallel-axis slabs make it exact)."""
    sc = BVH_SCENES[name]()
    rays = hit_test_rays(sc, name, 40000)
    engine.upload(sc)
    engine.build_bvh()
    prim, tuv = engine.closest_hit(rays)
    prim, tuv = prim.cpu().numpy(), tuv.cpu().numpy()
    os_ = oracle.OracleScene(sc)
    oprim, otuv = os_.closest_hit(rays)
    same = prim == oprim
    print(f"[hit] {name}: {len(rays)} rays, match {same.mean():.6f}, hits {np.mean(oprim >= 0):.3f}", flush=True)
    bad = np.where(~same)[0]
    assert len(bad) == 0, [(int(k), int(prim[k]), int(oprim[k]), tuv[k].tolist(), otuv[k].tolist()) for k in bad[:8]]
    np.testing.assert_array_equal(tuv, otuv)


@gpu
@pytest.mark.parametrize("name", ["box_room", "terrain_spheres", "soup_5000", "transformed"])
def test_near_axis_rays_never_miss_a_reference_hit(oracle, engine, name):
    """Nearly axis-parallel rays (direction components of 1e-3 .. 1e-37): the device walk never
    misses a hit the reference walk finds nor returns a farther one; closer hits, in boxes the
    reference's division-form slab test rounds away, stay under 1 in 1000 rays."""
    sc = BVH_SCENES[name]()
    rays = near_axis_test_rays(name, 16000)
    engine.upload(sc)
    engine.build_bvh()
    prim, tuv = engine.closest_hit(rays)
    prim, tuv = prim.cpu().numpy(), tuv.cpu().numpy()
    oprim, otuv = oracle.OracleScene(sc).closest_hit(rays)
    missed = np.where((oprim >= 0) & ((prim < 0) | (tuv[:, 0] > otuv[:, 0])))[0]
    assert len(missed) == 0, [(int(k), int(prim[k]), int(oprim[k]), tuv[k].tolist(), otuv[k].tolist()) for k in missed[:8]]
    same = prim == oprim
    np.testing.assert_array_equal(tuv[same], otuv[same])
    print(f"[near-axis] {name}: {len(rays)} rays, {int((~same).sum())} closer hits than the reference walk", flush=True)
    assert 1.0 - same.mean() <= 1e-3


@gpu
def test_render_transformed_meshes_radiance(oracle, engine):
    sc = scenes.transformed_scene()
    engine.upload(sc)
    engine.build_bvh()
    img = engine.render(256, 256, 16, 5, rr_start=3)[0].cpu().numpy()
    ref, _ = oracle.OracleScene(sc).render(256, 256, 16, 5, rr_start=3, threads=8)
    err = rel_rmse(img, ref)
    print(f"[render] transformed 256x256x16 d5 rr3 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_render_box_room_radiance(oracle, engine):
    sc = scenes.box_room()
    w = h = 64
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(w, h, frames=16, max_depth=5)
    img = final.cpu().numpy()
    ref, _ = oracle.OracleScene(sc).render(w, h, 16, 5, threads=8)
    err = rel_rmse(img, ref)
    print(f"[render] box_room 64x64x16 d5 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_render_terrain_spheres_radiance_with_rr(oracle, engine):
    sc = scenes.scaled_million(scale_cells=48, spheres=8, sphere_level=3, width=96, height=54)
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(96, 54, frames=4, max_depth=8, rr_start=3)
    img = final.cpu().numpy()
    ref, _ = oracle.OracleScene(sc).render(96, 54, 4, 8, rr_start=3, threads=8)
    err = rel_rmse(img, ref)
    print(f"[render] terrain_spheres 96x54x4 d8 rr3 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3


@gpu
def test_render_frame_offset_and_vertical_invert(oracle, engine):
    sc = scenes.box_room()
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(48, 40, frames=3, max_depth=3, frame_start=5, max_samples_offset=17, vertical_invert=True)
    ref, _ = oracle.OracleScene(sc).render(48, 40, 3, 3, frame_start=5, max_samples_offset=17, vertical_invert=True, threads=8)
    assert rel_rmse(final.cpu().numpy(), ref) <= 1e-3


@gpu
def test_render_gpu_raygen_index_policy(oracle, engine):
    from axomae_amd.nova_hip import INDEX_OPTIX
    sc = scenes.box_room()
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(48, 40, frames=4, max_depth=3, frame_start=3, max_samples_offset=64, index_policy=INDEX_OPTIX)
    oracle.set_index_policy(1)
    try:
        ref, _ = oracle.OracleScene(sc).render(48, 40, 4, 3, frame_start=3, max_samples_offset=64, threads=8)
    finally:
        oracle.set_index_policy(0)
    assert rel_rmse(final.cpu().numpy(), ref) <= 1e-3


@gpu
@pytest.mark.parametrize("scene_name", ["box_room", "terrain_spheres"])
def test_render_passes_match_oracle(oracle, engine, scene_name):
    """Normal, depth and emissive passes (integrator_flags) against the oracle's, pixel by
    pixel: the GPU walk settles ties like the reference, so the values are identical."""
    from axomae_amd import nova_hip as nh
    sc = BVH_SCENES[scene_name]()
    engine.upload(sc)
    engine.build_bvh()
    osc = oracle.OracleScene(sc)
    w, h = 72, 40
    for pass_, flag in [(1, nh.INTEGRATOR_NORMAL), (2, nh.INTEGRATOR_DEPTH), (3, nh.INTEGRATOR_EMISSIVE)]:
        final, _ = engine.render(w, h, frames=3, max_depth=4, frame_start=1, integrator_flags=nh.INTEGRATOR_PATH | flag, z_near=0.1, z_far=50.0)
        ref, _ = osc.render_pass(w, h, 3, pass_, z_near=0.1, z_far=50.0, frame_start=1)
        img = final.cpu().numpy()
        same = np.all(img == ref, axis=2).mean()
        print(f"[passes] {scene_name} pass {pass_}: identical pixels {same:.5f}", flush=True)
        assert same >= 0.999


@gpu
def test_render_integrator_flag_errors(engine):
    """integrator_dispatch: a flag word without PATH is INVALID_INTEGRATOR, PATH without a
    render pass INVALID_RENDER_MODE; both surface as RENDERER_EXCEPTION (16)."""
    from axomae_amd import nova_hip as nh
    sc = scenes.box_room()
    engine.upload(sc)
    engine.build_bvh()
    for flags in (nh.INTEGRATOR_NORMAL, nh.INTEGRATOR_PATH):
        with pytest.raises(nh.NovaHipError) as err:
            engine.render(16, 16, frames=1, max_depth=1, integrator_flags=flags)
        assert err.value.code == 16
    a, _ = engine.render(16, 16, frames=1, max_depth=1, integrator_flags=nh.INTEGRATOR_PATH | nh.INTEGRATOR_COMBINED)
    b, _ = engine.render(16, 16, frames=1, max_depth=1)
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


@gpu
def test_brdf_lut_bake_matches_oracle(oracle):
    """The split-sum BRDF table (brdf_lookup_table_baker.frag) as a bake job beside an
    irradiance job in the same launch."""
    from axomae_amd import nova_hip as nh
    env = scenes.sky_envmap(64, 32, (0.3, 0.9, 0.2))
    lut, irr = nh.bake_batch([(None, 48, 40, nh.BAKE_BRDF_LUT, 0.0), (env, 16, 8, nh.BAKE_IRRADIANCE, 0.0)], 1024)
    ref = oracle.bake_brdf_lut(48, 40, 1024)
    print(f"[brdf-lut] max abs diff {np.abs(lut - ref).max():.3e}", flush=True)
    np.testing.assert_allclose(lut, ref, rtol=1e-5, atol=1e-6)
    assert rel_rmse(irr, oracle.bake_irradiance(env, 16, 8, 1024)) <= 1e-3


@gpu
def test_bakes_match_cpu_bake(oracle):
    from axomae_amd import nova_hip
    maps = [scenes.sky_envmap(128, 64, (0.2, 0.9, 0.4), sun_intensity=20.0, sun_size=0.01), scenes.sky_envmap(256, 128, (-0.5, 0.3, 0.8))]
    sizes = [(16, 8), (32, 16)]
    gpu_irr = nova_hip.bake("irradiance", maps, sizes, 128)
    for m, s, g in zip(maps, sizes, gpu_irr):
        ref = oracle.bake_irradiance(m, s[0], s[1], 128)
        assert rel_rmse(g, ref) <= 1e-3
    rough = [0.0, 0.5]
    gpu_pf = nova_hip.bake("prefilter", maps, sizes, 256, roughness=rough)
    for m, s, g, r in zip(maps, sizes, gpu_pf, rough):
        ref = oracle.bake_prefilter(m, s[0], s[1], r, 256)
        assert rel_rmse(g, ref) <= 1e-3


@gpu
def test_render_fp32_material_textures(oracle, engine):
    from test_device_code_on_cpu import fp32_textured_scene
    sc = fp32_textured_scene()
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(32, 20, frames=2, max_depth=4)
    ref, _ = oracle.OracleScene(sc).render(32, 20, 2, 4, threads=8)
    assert rel_rmse(final.cpu().numpy(), ref) <= 1e-3


@gpu
def test_bake_batch_one_launch_matches_cpu_bake(oracle):
    """Irradiance (pgc3d and Sobol samplers) and a prefilter mip chain of several maps in a
    single nova_hip_bake_batch launch, each output against the CPU bake."""
    from axomae_amd import nova_hip
    maps = [scenes.sky_envmap(256, 128, (0.2, 0.9, 0.4), sun_intensity=30.0, sun_size=0.005), scenes.sky_envmap(128, 64, (-0.5, 0.3, 0.8))]
    jobs, refs = [], []
    for m in maps:
        jobs.append((m, 32, 16, nova_hip.BAKE_IRRADIANCE, 0.0))
        refs.append(lambda m=m: oracle.bake_irradiance(m, 32, 16, 256))
        jobs.append((m, 32, 16, nova_hip.BAKE_IRRADIANCE_SOBOL, 0.0))
        refs.append(lambda m=m: oracle.bake_irradiance(m, 32, 16, 256, sobol=True))
        for i in range(6):
            w, h, r = max(64 >> i, 1), max(32 >> i, 1), i / 5.0
            jobs.append((m, w, h, nova_hip.BAKE_PREFILTER, r))
            refs.append(lambda m=m, w=w, h=h, r=r: oracle.bake_prefilter(m, w, h, r, 256))
    outs = nova_hip.bake_batch(jobs, 256)
    for k, (o, ref) in enumerate(zip(outs, refs)):
        err = rel_rmse(o, ref())
        assert err <= 1e-3, (k, jobs[k][1:], err)


@gpu
@pytest.mark.parametrize("name", ["soup_200k", "million_triangles"])
def test_bvh_primitive_order_bit_exact_large(oracle, engine, name):
    """Scenes whose upper levels run the many-workgroup node path (nodes above 16,384
    primitives): the same prim_idx, node count, leaves and boxes as the reference builder."""
    sc = soup_scene(200000, 21, lo=-20.0, hi=20.0, size=0.2) if name == "soup_200k" else scenes.million_triangles()
    engine.upload(sc)
    info = engine.build_bvh()
    print(f"[bvh] {name}: {sc.triangle_count} triangles, build {info.build_ms:.1f} ms", flush=True)
    prim, nodes = engine.bvh_arrays()
    oprim, onodes = oracle.OracleScene(sc).bvh_arrays()
    np.testing.assert_array_equal(prim, oprim)
    assert info.binary_nodes == len(onodes)
    assert _leaf_set(nodes) == _leaf_set(onodes)
    key = lambda arr: sorted(tuple(np.concatenate([a["min"], a["max"]]).tolist()) + (int(a["count"]),) for a in arr)
    assert key(nodes) == key(onodes)


@gpu
@pytest.mark.parametrize("frames", [8, 16, 64])
def test_wave_coherent_camera_walk_matches_the_per_ray_walk(engine, frames):
    """Camera rays of pixel-major batches (>= 8 frames) are traced by one walk per wave of 64
    consecutive rays (the last packet partly empty at 50x37); on these scenes the hits, and so
    the images and accumulators, equal the per-ray walk's bit for bit. (The two walks can
    differ on a ray grazing the edge two triangles share, where Moller-Trumbore's t lies
    before the triangle's box and the result depends on which leaf is reached first.)"""
    from axomae_amd.nova_hip import RENDER_COUNT_TRAVERSAL, RENDER_PER_RAY_TRAVERSAL

    for sc in (scenes.box_room(), scenes.scaled_million(scale_cells=48, spheres=8, sphere_level=3, width=96, height=54)):
        engine.upload(sc)
        engine.build_bvh()
        for w, h in ((48, 40), (50, 37)):
            print(f"[coherent] {sc.name} {w}x{h} frames {frames}", flush=True)
            a, acc_a = engine.render(w, h, frames=frames, max_depth=4, rr_start=2)
            b, acc_b = engine.render(w, h, frames=frames, max_depth=4, rr_start=2, flags=RENDER_PER_RAY_TRAVERSAL)
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
            np.testing.assert_array_equal(acc_a.cpu().numpy(), acc_b.cpu().numpy())
            # the instrumented render takes the same walks
            c, _ = engine.render(w, h, frames=frames, max_depth=4, rr_start=2, flags=RENDER_COUNT_TRAVERSAL)
            np.testing.assert_array_equal(a.cpu().numpy(), c.cpu().numpy())


@gpu
@pytest.mark.parametrize("frames,depth,rr", [(1, 8, 3), (8, 5, -1), (16, 6, 2)])
def test_last_bounces_in_one_launch_match_per_bounce_launches(engine, frames, depth, rr):
    """Once few paths remain, the last bounces of a batch run in one launch, one path per lane
    (tail_kernel); images, accumulators and traced segments equal those of one extend and
    one shade launch per bounce (NOVA_HIP_RENDER_NO_TAIL) bit for bit."""
    from axomae_amd.nova_hip import RENDER_COUNT_TRAVERSAL, RENDER_NO_TAIL

    for sc in (scenes.box_room(), scenes.transformed_scene(), scenes.scaled_million(scale_cells=48, spheres=8, sphere_level=3, width=96, height=54)):
        engine.upload(sc)
        engine.build_bvh()
        a, acc_a = engine.render(64, 48, frames=frames, max_depth=depth, rr_start=rr, flags=RENDER_COUNT_TRAVERSAL)
        sa = engine.render_stats()
        b, acc_b = engine.render(64, 48, frames=frames, max_depth=depth, rr_start=rr, flags=RENDER_COUNT_TRAVERSAL | RENDER_NO_TAIL)
        sb = engine.render_stats()
        print(f"[tail] {sc.name} frames {frames} depth {depth} rr {rr}: segments {sa.segments} (tail {sa.tail_segments}) vs {sb.segments}", flush=True)
        assert sa.tail_segments > 0 and sb.tail_segments == 0
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
        np.testing.assert_array_equal(acc_a.cpu().numpy(), acc_b.cpu().numpy())
        assert (sa.segments, sa.node_visits, sa.triangle_tests, sa.escaped) == (sb.segments, sb.node_visits, sb.triangle_tests, sb.escaped)


@gpu
def test_render_of_loaded_assets(oracle, engine, tmp_path):
    """A scene read from .obj/.mtl (box faces, two materials, one emissive) with a Radiance
    sky renders on the device like the oracle."""
    from axomae_amd import loaders

    (tmp_path / "m.mtl").write_text("newmtl white\nKd 0.7 0.7 0.7\nPr 0.6\n\nnewmtl lamp\nKd 0.2 0.2 0.2\nKe 4 3.5 3\n\nnewmtl glass\nKd 1 1 1\nPr 0.05\nd 0.2\nNi 1.5\n")
    lines = ["mtllib m.mtl"]
    vs = [(-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2), (-0.5, 1.5, -0.5), (0.5, 1.5, -0.5), (0.5, 1.5, 0.5), (-0.5, 1.5, 0.5),
          (-0.6, 0.2, 0.6), (0.4, 0.2, 0.9), (0.0, 0.9, 0.7)]
    lines += [f"v {x} {y} {z}" for x, y, z in vs]
    lines += ["vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1", "vn 0 1 0", "vn 0 -1 0"]
    lines += ["usemtl white", "f 1/1/1 4/4/1 3/3/1 2/2/1", "usemtl lamp", "f 5/1/2 6/2/2 7/3/2 8/4/2", "usemtl glass", "f 9 10 11"]
    obj = tmp_path / "s.obj"
    obj.write_text("\n".join(lines) + "\n")
    w, h = 64, 32
    rgbe = np.zeros((h, w, 4), dtype=np.uint8)
    rgbe[..., :3] = [120, 160, 230]
    rgbe[..., 3] = 128  # 2^-8 scale: sky ~0.5 .. 0.9
    hdr = tmp_path / "sky.hdr"
    hdr.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + rgbe.tobytes())
    sc = scenes.Scene(name="loaded")
    loaders.load_obj(str(obj), sc)
    loaders.load_envmap(str(hdr), sc)
    scenes.set_camera(sc, (0.0, 1.2, 4.0), (0.0, 0.5, 0.0), 45.0, 1.5)
    engine.upload(sc)
    engine.build_bvh()
    final, _ = engine.render(48, 32, frames=8, max_depth=4)
    ref, _ = oracle.OracleScene(sc).render(48, 32, 8, 4, threads=8)
    err = rel_rmse(final.cpu().numpy(), ref)
    print(f"[render] loaded .obj/.hdr scene 48x32x8 d4 rel_rmse={err:.3e}", flush=True)
    assert err <= 1e-3
