// Host build of the device path stages (test infrastructure): the scene flattening,
// camera ray, Moller-Trumbore and per-segment shading functions the kernels run, driven
// by a CPU loop with a brute-force closest hit. It checks the device code's arithmetic
// and control flow against the oracle on machines without a GPU.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../axomae_amd/csrc/device_path.h"
#include "../../axomae_amd/csrc/device_scene.h"
#include "../../axomae_amd/csrc/device_trace.h"

using namespace nova_hip;

namespace {
const uint32_t kTable[256][32] = {
#include "../../axomae_amd/csrc/sobol_table.inc"
};
}

namespace {
// Scene flattening as nova_hip_scene_upload does it, on the host.
struct EmuScene {
  uint64_t ntri = 0;
  std::vector<TriMT> tri;
  std::vector<ShadeTri> shade;
  std::vector<float> aabb, cent;
  std::vector<MeshXf> mxf;
  std::vector<float4> xf_inv;
};

void flatten_scene(const nova_hip_scene_desc *sd, EmuScene &es) {
  uint64_t nvert = 0, ntri = 0;
  for (uint32_t m = 0; m < sd->mesh_count; m++) {
    nvert += sd->meshes[m].vertex_count;
    ntri += sd->meshes[m].index_count / 3;
  }
  std::vector<float> pos(nvert * 3), nrm(nvert * 3), tan(nvert * 3), bit(nvert * 3), uv(nvert * 2), xf(sd->mesh_count * 16);
  std::vector<int32_t> mesh_mat(sd->mesh_count);
  std::vector<MeshXf> mxf(sd->mesh_count);
  std::vector<TriRef> refs(ntri);
  es.xf_inv.assign(sd->mesh_count * 3 + 3, make_float4(0.f, 0.f, 0.f, 0.f));
  uint64_t vo = 0, to = 0;
  for (uint32_t m = 0; m < sd->mesh_count; m++) {
    const nova_hip_mesh_desc &md = sd->meshes[m];
    std::memcpy(&pos[vo * 3], md.vertices, md.vertex_count * 12);
    std::memcpy(&nrm[vo * 3], md.normals, md.vertex_count * 12);
    std::memcpy(&tan[vo * 3], md.tangents, md.vertex_count * 12);
    std::memcpy(&bit[vo * 3], md.bitangents, md.vertex_count * 12);
    std::memcpy(&uv[vo * 2], md.uv, md.vertex_count * 8);
    std::memcpy(&xf[m * 16], md.transform, 64);
    mesh_xf_blocks(md.transform, mxf[m]);
    mesh_inverse_rows(md.transform, &es.xf_inv[m * 3]);
    mesh_mat[m] = md.material;
    for (uint64_t i = 0; i < md.index_count; i += 3)
      refs[to++] = TriRef{(uint32_t)(md.indices[i] + vo), (uint32_t)(md.indices[i + 1] + vo), (uint32_t)(md.indices[i + 2] + vo), m};
    vo += md.vertex_count;
  }
  es.ntri = ntri;
  es.tri.resize(ntri);
  es.shade.resize(ntri);
  es.aabb.resize(ntri * 6);
  es.cent.resize(ntri * 3);
  for (uint32_t t = 0; t < ntri; t++)
    flatten_triangle(t, refs[t], pos.data(), nrm.data(), tan.data(), bit.data(), uv.data(), xf.data(), mesh_mat.data(), es.tri.data(),
                     es.shade.data(), es.aabb.data(), es.cent.data());
  es.mxf = mxf;
}

struct HostStack {
  std::vector<uint32_t> v;
  void push(int sp, uint32_t x) {
    if ((size_t)sp >= v.size()) v.resize(sp + 64);
    v[sp] = x;
  }
  uint32_t at(int sp) const { return v[sp]; }
};
}  // namespace

// Closest hit through the device traversal code on a binary BVH given by the caller (the
// reference builder's output, which the GPU build reproduces): parent/depth/leaf maps and
// the 4-wide collapse are derived here as the GPU build derives them.
extern "C" int emu_closest_hit(const nova_hip_scene_desc *sd, const BNode *bnodes, int32_t nbnodes, const int32_t *prim_idx, const float *rays,
                               int64_t nrays, int32_t *prim_out, float *tuv_out, uint64_t *counts) {
  EmuScene es;
  flatten_scene(sd, es);
  const uint32_t n = (uint32_t)es.ntri;
  std::vector<int32_t> parent(nbnodes, -1), depth(nbnodes, 0), leaf_of_pos(n, -1);
  std::vector<int32_t> order{0};
  for (size_t k = 0; k < order.size(); k++) {
    const int32_t i = order[k];
    const BNode &b = bnodes[i];
    if (b.count > 0) {
      for (int32_t j = 0; j < b.count; j++) leaf_of_pos[b.left + j] = i;
      continue;
    }
    for (int32_t c = b.left; c <= b.left + 1; c++) {
      parent[c] = i;
      depth[c] = depth[i] + 1;
      order.push_back(c);
    }
  }
  std::vector<Node4> wide(1);
  std::vector<std::pair<int32_t, int32_t>> frontier{{0, 0}};
  for (size_t k = 0; k < frontier.size(); k++) {
    int32_t ch[4];
    const int m = collapse_children(bnodes, frontier[k].first, ch);
    Node4 nd;
    set_node_boxes(nd, bnodes, ch, m);
    for (int c = 0; c < 4; c++) {
      if (c >= m) {
        nd.child[c] = -1;
        continue;
      }
      const BNode &b = bnodes[ch[c]];
      if (b.count > 0) {
        nd.child[c] = leaf_ref(b.left);
      } else {
        nd.child[c] = (int32_t)wide.size();
        frontier.push_back({ch[c], (int32_t)wide.size()});
        wide.emplace_back();
      }
    }
    wide[frontier[k].second] = nd;
  }
  std::vector<TriMT> tri_leaf(n);
  for (uint32_t i = 0; i < n; i++) tri_leaf[i] = es.tri[prim_idx[i]];
  for (int32_t i = 0; i < nbnodes; i++)
    if (bnodes[i].count > 0) tri_leaf[bnodes[i].left].b.w = __builtin_bit_cast(float, bnodes[i].count);
  const TieTables tie{bnodes, parent.data(), depth.data(), leaf_of_pos.data(), es.xf_inv.data()};
  TraceCtx tc{wide.data(), tri_leaf.data(), nullptr, 0, &tie};
  HostStack st;
  uint32_t nv = 0, nt = 0;
  for (int64_t r = 0; r < nrays; r++) {
    const float *ry = rays + r * 6;
    float t, u, v;
    const int p = trace_closest<true>(tc, st, mk3(ry[0], ry[1], ry[2]), mk3(ry[3], ry[4], ry[5]), 0.000001f, t, u, v, nv, nt);
    prim_out[r] = p;
    tuv_out[r * 3] = p >= 0 ? t : 0.f;
    tuv_out[r * 3 + 1] = u;
    tuv_out[r * 3 + 2] = v;
  }
  if (counts) {
    counts[0] = nv;
    counts[1] = nt;
  }
  return 0;
}

// Renders (adj_grad == nullptr: accumulating into accum/final_out) or, for the adjoint,
// renders each frame and replays it with the gradient pass of shade_path.
static int emu_run(const nova_hip_scene_desc *sd, const nova_hip_render_params *p, float *accum, float *final_out, const float *adj_dimg,
                   float adj_scale, float *adj_grad) {
  EmuScene es;
  flatten_scene(sd, es);
  const uint64_t ntri = es.ntri;
  const std::vector<TriMT> &tri = es.tri;
  const std::vector<ShadeTri> &shade = es.shade;
  std::vector<MeshXf> mxf = es.mxf;
  std::vector<TexRec> texrec(sd->texture_count);
  std::vector<uint32_t> u32pool;
  std::vector<float> f32pool;
  for (uint32_t t = 0; t < sd->texture_count; t++) {
    const nova_hip_texture_desc &td = sd->textures[t];
    TexRec &r = texrec[t];
    r.width = td.width;
    r.height = td.height;
    r.channels = td.channels;
    r.invert_u = td.invert_u;
    r.invert_v = td.invert_v;
    r.format = td.format;
    if (td.format == NOVA_HIP_TEX_U32) {
      r.offset = u32pool.size();
      u32pool.insert(u32pool.end(), (const uint32_t *)td.data, (const uint32_t *)td.data + (size_t)td.width * td.height);
    } else {
      r.offset = f32pool.size();
      f32pool.insert(f32pool.end(), (const float *)td.data, (const float *)td.data + (size_t)td.width * td.height * td.channels);
    }
  }
  std::vector<MatRec> mats(sd->material_count);
  for (uint32_t m = 0; m < sd->material_count; m++) {
    const nova_hip_material_desc &d = sd->materials[m];
    MatRec &r = mats[m];
    r.albedo = d.albedo;
    r.metallic = d.metallic;
    r.normal = d.normal;
    r.emissive = d.emissive;
    r.albedo_f32 = d.albedo_f32;
    r.roughness_f32 = d.roughness_f32;
    for (int k = 0; k < 3; k++) {
      r.eta[k] = d.eta[k];
      r.k[k] = d.k[k];
    }
    r.anisotropy = d.anisotropy;
    r.emission_scale = d.emission_scale;
  }
  std::vector<float> lut(kAcosLut);
  for (int i = 0; i < kAcosLut; i++) lut[i] = (float)std::acos(((double)i / (double)kAcosLut) * 2.0 - 1.0);
  ShadeCtx sc{texrec.data(), u32pool.data(), f32pool.data(), mats.data(), mxf.data(), shade.data(), lut.data(), sd->envmap_texture,
              p->sampling_roughness_texture};

  const uint32_t npix = (uint32_t)p->width * (uint32_t)p->height;
  std::vector<float> partial((size_t)npix * 4);
  const TieTables xf_tab{nullptr, nullptr, nullptr, nullptr, es.xf_inv.data()};
  auto brute_hit = [&](f3 o, f3 d, float &tb, float &bu, float &bv) {
    tb = 3.402823466e+38f;
    bu = bv = 0.f;
    int best = -1;
    for (uint32_t t = 0; t < ntri; t++) {
      float tt, uu, vv;
      if (tri_hit(tri[t], [&]() { return &xf_tab; }, o, d, 0.000001f, tb, tt, uu, vv)) {
        tb = tt;
        bu = uu;
        bv = vv;
        best = (int)t;
      }
    }
    return best;
  };

  // ---- render passes: one unjittered ray per pixel, the same value every frame
  const int32_t fl = p->integrator_flags;
  const int pass_id = (fl & NOVA_HIP_INTEGRATOR_EMISSIVE) ? PASS_EMISSIVE
                      : (fl & NOVA_HIP_INTEGRATOR_DEPTH)  ? PASS_DEPTH
                      : (fl & NOVA_HIP_INTEGRATOR_NORMAL) ? PASS_NORMAL
                                                          : 0;
  if (pass_id != 0 && !adj_grad) {
    FrameParams fp;
    fp.width = p->width;
    fp.height = p->height;
    fp.frame0 = p->frame_start;
    fp.frame_minor = 0;
    fp.tiled = 0;
    fp.frames = p->frame_count;
    fp.offset = 0;
    fp.vinvert = p->vertical_invert;
    fp.policy = NOVA_HIP_INDEX_CPU;
    std::memcpy(fp.inv_P, p->inv_P, 64);
    std::memcpy(fp.inv_V, p->inv_V, 64);
    for (uint32_t px = 0; px < npix; px++) {
      f3 o, d;
      camera_ray_pass(fp, px, o, d);
      float tb, bu, bv;
      const int best = brute_hit(o, d, tb, bu, bv);
      const float4 v = pass_value(pass_id, fp, sc, &kTable[0][0], px, o, d, best, tb, bu, bv, p->z_near, p->z_far);
      float *dst = &partial[pixel_offset(fp, px)];
      dst[0] = v.x;
      dst[1] = v.y;
      dst[2] = v.z;
      dst[3] = v.w;
    }
    for (int f = p->frame_start; f < p->frame_start + p->frame_count; f++)
      for (size_t i = 0; i < (size_t)npix * 4; i++) {
        accum[i] += partial[i];
        final_out[i] = accum[i] / float(f + 1);
      }
    return 0;
  }

  // ---- frames: camera ray, then closest hit + shade per bounce, then accumulate
  for (int f = p->frame_start; f < p->frame_start + p->frame_count; f++) {
    FrameParams fp;
    fp.width = p->width;
    fp.height = p->height;
    fp.frame0 = f;
    fp.frame_minor = 0;
    fp.tiled = 0;
    fp.frames = 1;
    fp.offset = p->max_samples_offset;
    fp.vinvert = p->vertical_invert;
    fp.policy = p->index_policy;
    std::memcpy(fp.inv_P, p->inv_P, 64);
    std::memcpy(fp.inv_V, p->inv_V, 64);
    const AdjointCtx adj{partial.data(), adj_dimg, adj_scale, adj_grad, nullptr, 0u, 0u};
    for (int pass = 0; pass < (adj_grad ? 2 : 1); pass++)
    for (uint32_t slot = 0; slot < npix; slot++) {
      f3 o, d;
      camera_ray(fp, &kTable[0][0], slot, o, d);
      f3 T = mk3(1, 1, 1), L = mk3(0, 0, 0);
      float A = 0.f;
      uint32_t dim = 2;
      AdjointCtx padj = adj;
      padj.ltot = partial.data() + pixel_offset(fp, slot);  // this path's primal radiance
      for (int k = 0; k <= p->max_depth; k++) {
        float tb, bu, bv;
        const int best = brute_hit(o, d, tb, bu, bv);
        ShadeParams sp;
        sp.fp = fp;
        sp.bounce = k;
        sp.last = k == p->max_depth;
        sp.rr = (p->rr_start >= 0 && k >= p->rr_start) ? 1 : 0;
        sp.dim0 = dim;
        f3 no, nd;
        bool esc;
        bool cont = pass == 0 ? shade_path(sp, sc, &kTable[0][0], slot, o, d, best, tb, bu, bv, T, L, A, no, nd, esc)
                              : shade_path<true>(sp, sc, &kTable[0][0], slot, o, d, best, tb, bu, bv, T, L, A, no, nd, esc, &padj);
        dim += 4 + (sp.rr ? 1 : 0);
        if (!cont) break;
        o = no;
        d = nd;
      }
      if (pass == 1) continue;
      float *dst = &partial[pixel_offset(fp, slot)];
      dst[0] = L.x;
      dst[1] = L.y;
      dst[2] = L.z;
      dst[3] = A;
    }
    if (adj_grad) continue;
    for (size_t i = 0; i < (size_t)npix * 4; i++) {
      accum[i] += partial[i];
      final_out[i] = accum[i] / float(f + 1);
    }
  }
  return 0;
}

extern "C" int emu_render(const nova_hip_scene_desc *sd, const nova_hip_render_params *p, float *accum, float *final_out) {
  return emu_run(sd, p, accum, final_out, nullptr, 0.f, nullptr);
}

// Texel gradients (FP32 pool layout) from the device path's replay code run on the host.
extern "C" int emu_adjoint(const nova_hip_scene_desc *sd, const nova_hip_render_params *p, const float *d_image, float scale, float *grad) {
  return emu_run(sd, p, nullptr, nullptr, d_image, scale, grad);
}

// Sampler slot of every path id of a width x height x frames batch in the renderers' path
// numbering (pixel-major, frame-minor; pixels in 8x8 tiles when tiled), out: npix * frames.
extern "C" int emu_path_slots(int32_t width, int32_t height, int32_t frames, int32_t tiled, uint32_t *out) {
  FrameParams fp{};
  fp.width = width;
  fp.height = height;
  fp.frames = frames;
  fp.frame_minor = 1;
  fp.tiled = tiled;
  const uint32_t n = (uint32_t)width * (uint32_t)height * (uint32_t)frames;
  for (uint32_t pid = 0; pid < n; pid++) out[pid] = path_slot(fp, pid);
  return 0;
}

// Sobol samples of dims [dim0, dim0+5) for the given indices, from the direction table and
// from the shade kernels' per-nibble tables (out_direct, out_nibble: n x 5 floats).
extern "C" int emu_sobol_nibble(uint32_t dim0, const uint32_t *idx, int n, float *out_direct, float *out_nibble) {
  uint32_t tab[5 * 128];
  sobol_nibble_fill(&kTable[0][0], dim0, 5, tab, 0, 1);
  const SobolNibbles s{tab, dim0};
  for (int i = 0; i < n; i++) {
    const uint32_t g = idx[i] ^ (idx[i] >> 1);
    for (uint32_t d = 0; d < 5; d++) {
      if (dim0 + d >= 256) continue;
      out_direct[i * 5 + d] = sobol_dim(&kTable[0][0], dim0 + d, g);
      out_nibble[i * 5 + d] = sobol_dim(s, dim0 + d, g);
    }
  }
  return 0;
}
