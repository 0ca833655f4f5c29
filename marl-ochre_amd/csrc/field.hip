// field.hip — field kernels for gfx950: masked 7-point stencil, boundary
// sources, Beer-Lambert light scan, AoS<->SoA transposes, occupancy packing.
//
// Reference semantics:
//   stencil   crates/marl-field/src/field.rs:107-203 (and the WGSL statement
//             crates/marl-gpu/src/shaders/field_diffuse.wgsl:55-95)
//   sources   crates/marl-field/src/field.rs:231-250
//   light     crates/marl-field/src/light.rs:29-59
#include "marl_internal.h"

// ---------------------------------------------------------------------------
// K1: masked stencil.
//
// A block of 256 threads (4 wave64) owns a 64 (x) x 4*RY (y) tile and marches
// it through a z-chunk of zc planes.  Each thread keeps, for each of its RY
// rows, the 12 species of planes z-1, z, z+1 (and a prefetch of z+2) in
// registers; x/y neighbours come from an LDS copy of plane z with a one-voxel
// halo.  HBM traffic per substep is one read and one write of every species
// plane plus the 1-bit mask, plus two halo planes per z-chunk.  The niche
// factor is computed in registers from species 7 of the centre voxel.
//
// The per-voxel arithmetic is the reference's, in its order and without
// contraction, so the result is bit-identical to the f32 CPU statement.

const TileShape kTiles[] = {
    {1, 8}, {1, 16}, {1, 32}, {1, 64}, {2, 8}, {2, 16}, {2, 32}, {2, 64}, {4, 16}, {4, 32},
};
const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

int marl_default_tile(const marl_dims& d) {
  (void)d;
  return 2;  // RY=1, zc=32 until the autotuner says otherwise
}

// Raw buffer accesses: the voxel's byte offset lives in one VGPR, the
// species-plane and z-plane offsets in SGPRs.  Accesses beyond the buffer's
// num_records read 0 and drop writes instead of faulting.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
#define MARL_RSRC_DW3 0x00020000
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, MARL_RSRC_DW3);
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}

// Load species s at plane zz; with SRC the boundary sources of field.rs:231-250
// are applied on the fly (the first substep of a tick reads the sourced field).
template <bool SRC>
__device__ __forceinline__ float ld_field(rsrc_t r, const StencilParams& p, int s, uint32_t Pb, uint32_t XYb,
                                          int Z, uint32_t voff, int zz) {
  float v = bload(r, voff, (uint32_t)s * Pb + (uint32_t)zz * XYb);
  if (SRC) {
    if (zz == 0) {
      if (s == 1) v = fminf(v + p.src_ox, p.c_max);
      if (s == 3) v = fminf(v + p.src_c, p.c_max);
    }
    if (zz == Z - 1 && s == 2) v = fminf(v + p.src_red, p.c_max);
  }
  return v;
}

template <int RY, bool SRC>
__global__ void __launch_bounds__(256) k_stencil(const float* __restrict__ src, float* __restrict__ dst,
                                                 const uint32_t* __restrict__ occ, int X, int Y, int Z, int XW,
                                                 int zc, int ntx, int nty, int ntz, StencilParams p) {
  constexpr int TY = 4 * RY;
  constexpr int LW = 66;
  constexpr int S = MARL_S_EXT;
  // one spare LDS row (TY + 2) and spare mask slots take the writes of lanes
  // that own no halo value, so those writes need no branch
  __shared__ float lds[S][TY + 3][LW];
  __shared__ uint32_t mlds[(TY + 2) * 4 + 256];

  // XCD-aware tile order: hardware dispatches block b to XCD b % 8; give each
  // XCD a contiguous run of logical tiles so y-neighbour tiles (which share
  // halo rows) sit in the same L2.
  const int nb = ntx * nty * ntz;
  int b = blockIdx.x;
  const int nb8 = nb & ~7;
  if (b < nb8) b = (b & 7) * (nb8 >> 3) + (b >> 3);
  const int tx = b % ntx;
  const int ty = (b / ntx) % nty;
  const int tz = b / (ntx * nty);

  const int tid = threadIdx.x;
  const int lx = tid & 63, wy = tid >> 6;
  const int x0 = tx * 64, y0 = ty * TY, z0 = tz * zc;
  const int z1 = min(z0 + zc, Z);
  const int x = x0 + lx;
  const bool xin = x < X;
  const uint32_t XY = (uint32_t)X * (uint32_t)Y;
  const uint32_t XYb = XY * 4u, Pb = XYb * (uint32_t)Z;
  const uint32_t XWb = (uint32_t)XW * 4u, MZb = XWb * (uint32_t)Y;
  const rsrc_t rs = make_rsrc(src, Pb * S);
  const rsrc_t rd = make_rsrc(dst, Pb * S);
  const rsrc_t rm = make_rsrc(occ, MZb * (uint32_t)Z);

  // Every load below is unconditional: lanes outside the grid (partial tiles)
  // read voxel 0 of the plane instead, a value they never use, so every access
  // stays inside the buffer and the loop body has no branches around loads.
  bool valid[RY];
  int yr[RY];
  uint32_t vo[RY], mo[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    yr[r] = y0 + wy + 4 * r;
    valid[r] = xin && yr[r] < Y;
    vo[r] = valid[r] ? ((uint32_t)yr[r] * (uint32_t)X + (uint32_t)x) * 4u : 0u;
    mo[r] = valid[r] ? (uint32_t)yr[r] * XWb + (uint32_t)(x >> 5) * 4u : 0u;
  }
  // halo source offset of this thread: rows for waves 0-1, columns for waves 2-3
  int hrow, hcol;
  uint32_t hvo;
  bool hw;  // lanes that own a halo slot
  if (wy < 2) {
    const int yy = (wy == 0) ? y0 - 1 : y0 + TY;
    hrow = (wy == 0) ? 0 : TY + 1;
    hcol = lx + 1;
    hw = yy >= 0 && yy < Y && xin;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)x) * 4u : 0u;
  } else {
    const int xx = (wy == 2) ? x0 - 1 : x0 + 64;
    const int yy = y0 + (lx < TY ? lx : 0);
    hrow = (lx < TY ? lx : 0) + 1;
    hcol = (wy == 2) ? 0 : LW - 1;
    hw = lx < TY && xx >= 0 && xx < X && yy < Y;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)xx) * 4u : 0u;
  }
  // mask words of plane z for rows y0-1 .. y0+TY, word columns x0/32-1 .. +2
  const int mrow = tid >> 2, mwc = tid & 3;
  const int myy = y0 - 1 + mrow, mwx = (x0 >> 5) - 1 + mwc;
  const bool mw = tid < (TY + 2) * 4;
  const bool mvalid = mw && myy >= 0 && myy < Y && mwx >= 0 && mwx < XW;
  const uint32_t mvo = mvalid ? (uint32_t)myy * XWb + (uint32_t)mwx * 4u : 0u;
  const int mslot = mw ? tid : (TY + 2) * 4 + tid;
  if (!hw) {  // park the halo write of lanes without a halo slot in the spare row
    hrow = TY + 2;
    hcol = lx;
  }

  auto ldq = [&](float (&q)[RY][S], uint32_t (&m)[RY], int zz) {
    // planes beyond the grid are never consumed; clamp the offset into range
    const int zl = zz < Z ? zz : Z - 1;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int s = 0; s < S; ++s) q[r][s] = ld_field<SRC>(rs, p, s, Pb, XYb, Z, vo[r], zl);
      m[r] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mo[r], (int)((uint32_t)zl * MZb), 0);
    }
  };

  float prv[RY][S], cur[RY][S], nxt[RY][S], nn[RY][S];
  uint32_t mp[RY], mc[RY], mn[RY], mnn[RY];
  ldq(cur, mc, z0);
  ldq(nxt, mn, z0 + 1);
  ldq(nn, mnn, z0 + 2);
  if (z0 > 0) {
    ldq(prv, mp, z0 - 1);
  } else {
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      mp[r] = 0u;
#pragma unroll
      for (int s = 0; s < S; ++s) prv[r][s] = 0.0f;
    }
  }

  for (int z = z0; z < z1; ++z) {
    // halo values and mask words of plane z (issued first, waited on first)
    float hv[S];
    const int zl = z;
#pragma unroll
    for (int s = 0; s < S; ++s) hv[s] = ld_field<SRC>(rs, p, s, Pb, XYb, Z, hvo, zl);
    const uint32_t mword = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mvo, (int)((uint32_t)zl * MZb), 0);
    // keep the halo loads ahead of the prefetch so the wait before the barrier
    // leaves the prefetch in flight
    __builtin_amdgcn_sched_barrier(0);
    // prefetch plane z+3
    float n3[RY][S];
    uint32_t mn3[RY];
    ldq(n3, mn3, z + 3);
    // plane z into LDS
#pragma unroll
    for (int r = 0; r < RY; ++r)
#pragma unroll
      for (int s = 0; s < S; ++s) lds[s][wy + 4 * r + 1][lx + 1] = cur[r][s];
#pragma unroll
    for (int s = 0; s < S; ++s) lds[s][hrow][hcol] = hv[s];
    mlds[mslot] = mvalid ? mword : 0u;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int row = wy + 4 * r + 1;
      const int y = yr[r];
      auto bitat = [&](int rw, int c) -> uint32_t {
        const int i = c + 32;
        return (mlds[rw * 4 + (i >> 5)] >> (i & 31)) & 1u;
      };
      const uint32_t zo = (uint32_t)z * XYb;
      // neighbour flags with bitwise operators: every mask read is unconditional
      const uint32_t bc = bitat(row, lx), bxm = bitat(row, lx - 1), bxp = bitat(row, lx + 1);
      const uint32_t bym = bitat(row - 1, lx), byp = bitat(row + 1, lx);
      const uint32_t bzm = (mp[r] >> (x & 31)) & 1u, bzp = (mn[r] >> (x & 31)) & 1u;
      const bool vxm = (uint32_t)(x > 0) & (bxm ^ 1u);
      const bool vxp = (uint32_t)(x < X - 1) & (bxp ^ 1u);
      const bool vym = (uint32_t)(y > 0) & (bym ^ 1u);
      const bool vyp = (uint32_t)(y < Y - 1) & (byp ^ 1u);
      const bool vzm = (uint32_t)(z > 0) & (bzm ^ 1u);
      const bool vzp = (uint32_t)(z < Z - 1) & (bzp ^ 1u);
      const float structural = cur[r][7];
      const float niche = 1.0f - p.alpha_eps * structural / (p.k_eps + structural);
      float out[S];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const float c = cur[r][s];
        const float dd = p.d_voxel[s] * niche;
        const float lxm = lds[s][row][lx];
        const float lxp = lds[s][row][lx + 2];
        const float lym = lds[s][row - 1][lx + 1];
        const float lyp = lds[s][row + 1][lx + 1];
        const float xm = vxm ? lxm : c;
        const float xp = vxp ? lxp : c;
        const float ym = vym ? lym : c;
        const float yp = vyp ? lyp : c;
        const float zm = vzm ? prv[r][s] : c;
        const float zp = vzp ? nxt[r][s] : c;
        const float lap = xm + xp + ym + yp + zm + zp - 6.0f * c;
        const float decay = p.lambda_decay[s] * c;
        const float nc = c + p.dt_sub * (dd * lap - decay);
        // occupied centre keeps c; a bit select keeps the compiler from
        // turning this into a branch around the neighbour reads
        const uint32_t keep = 0u - bc;
        out[s] = __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, c) & keep) |
                                              (__builtin_bit_cast(uint32_t, fmaxf(nc, 0.0f)) & ~keep));
      }
      if (valid[r]) {
#pragma unroll
        for (int s = 0; s < S; ++s) bstore(rd, out[s], vo[r], (uint32_t)s * Pb + zo);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        prv[r][s] = cur[r][s];
        cur[r][s] = nxt[r][s];
        nxt[r][s] = nn[r][s];
        nn[r][s] = n3[r][s];
      }
      mp[r] = mc[r];
      mc[r] = mn[r];
      mn[r] = mnn[r];
      mnn[r] = mn3[r];
    }
  }
}

template <int RY, bool SRC>
static hipError_t launch_stencil_t(marl_ctx* ctx, const StencilParams& p, int zc) {
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const int ntx = (X + 63) / 64, nty = (Y + 4 * RY - 1) / (4 * RY), ntz = (Z + zc - 1) / zc;
  const int nb = ntx * nty * ntz;
  const float* src = ctx->field[ctx->cur];
  float* dst = ctx->field[ctx->cur ^ 1];
  hipLaunchKernelGGL((k_stencil<RY, SRC>), dim3(nb), dim3(256), 0, ctx->stream, src, dst, ctx->occ, X, Y, Z,
                     ctx->xw, zc, ntx, nty, ntz, p);
  return hipGetLastError();
}

int marl_launch_substep(marl_ctx* ctx, const StencilParams& p, int tile) {
  if (tile < 0 || tile >= kNumTiles) tile = marl_default_tile(ctx->d);
  const TileShape t = kTiles[tile];
  const int zc = t.zc < ctx->d.z ? t.zc : ctx->d.z;
  hipError_t e = hipErrorInvalidValue;
  const bool src = p.apply_sources != 0;
  switch (t.ry) {
    case 1: e = src ? launch_stencil_t<1, true>(ctx, p, zc) : launch_stencil_t<1, false>(ctx, p, zc); break;
    case 2: e = src ? launch_stencil_t<2, true>(ctx, p, zc) : launch_stencil_t<2, false>(ctx, p, zc); break;
    case 4: e = src ? launch_stencil_t<4, true>(ctx, p, zc) : launch_stencil_t<4, false>(ctx, p, zc); break;
  }
  MARL_HIP(ctx, e);
  ctx->cur ^= 1;
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// Boundary sources in place (field.rs:231-250): one thread per (x, y).
__global__ void k_boundary(float* __restrict__ f, int X, int Y, int Z, float src_ox, float src_c, float src_red,
                           float c_max) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t XY = (size_t)X * Y, P = XY * Z;
  if (i >= (int)XY) return;
  f[1 * P + i] = fminf(f[1 * P + i] + src_ox, c_max);
  f[3 * P + i] = fminf(f[3 * P + i] + src_c, c_max);
  const size_t b = (size_t)(Z - 1) * XY + i;
  f[2 * P + b] = fminf(f[2 * P + b] + src_red, c_max);
}

int marl_launch_boundary(marl_ctx* ctx, const marl_sim_params* sim) {
  const int n = ctx->d.x * ctx->d.y;
  hipLaunchKernelGGL(k_boundary, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->field[ctx->cur], ctx->d.x,
                     ctx->d.y, ctx->d.z, sim->source_rate_oxidant, sim->source_rate_carbon,
                     sim->source_rate_reductant, sim->c_max);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// K2: Beer-Lambert light (light.rs:29-59).  One lane per (x, y) column,
// coalesced across x; the z loop carries the intensity in a register.  The
// sequential multiply order of the reference is kept per column.
__global__ void __launch_bounds__(64) k_light(const float* __restrict__ s4, const uint32_t* __restrict__ occ,
                                              float* __restrict__ light, int X, int Y, int Z, int XW, float I0,
                                              float e_cell, float nchem, float floor_) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= X) return;
  const size_t XY = (size_t)X * Y;
  const size_t col = (size_t)y * X + x;
  const uint32_t bit = 1u << (x & 31);
  const size_t wcol = (size_t)y * XW + (x >> 5);
  float I = I0;
  constexpr int U = 8;
  int z = 0;
  for (; z + U <= Z; z += U) {
    float a[U];
    uint32_t w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = s4[(size_t)(z + u) * XY + col];
      w[u] = occ[(size_t)(z + u) * Y * XW + wcol];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      light[(size_t)(z + u) * XY + col] = I;
      if (w[u] & bit) I *= e_cell;
      I *= marl_expf(nchem * a[u]);
      if (I < floor_) I = 0.0f;
    }
  }
  for (; z < Z; ++z) {
    light[(size_t)z * XY + col] = I;
    if (occ[(size_t)z * Y * XW + wcol] & bit) I *= e_cell;
    I *= marl_expf(nchem * s4[(size_t)z * XY + col]);
    if (I < floor_) I = 0.0f;
  }
}

int marl_launch_light(marl_ctx* ctx, const marl_sim_params* sim) {
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const float* s4 = ctx->field[ctx->cur] + (size_t)4 * ctx->nvox;
  const float e_cell = marl_expf(-sim->cell_absorption);
  hipLaunchKernelGGL(k_light, dim3((X + 63) / 64, Y), dim3(64), 0, ctx->stream, s4, ctx->occ, ctx->light, X, Y, Z,
                     ctx->xw, sim->surface_intensity, e_cell, -sim->chemical_absorption, sim->light_floor);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// Layout kernels.
__global__ void k_aos_to_soa(const float* __restrict__ aos, float* __restrict__ soa, size_t nvox) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const float4* a = reinterpret_cast<const float4*>(aos + v * MARL_S_EXT);
  const float4 q0 = a[0], q1 = a[1], q2 = a[2];
  const float vals[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) soa[(size_t)s * nvox + v] = vals[s];
}

__global__ void k_soa_to_aos(const float* __restrict__ soa, float* __restrict__ aos, size_t nvox) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  float vals[12];
#pragma unroll
  for (int s = 0; s < MARL_S_EXT; ++s) vals[s] = soa[(size_t)s * nvox + v];
  float4* a = reinterpret_cast<float4*>(aos + v * MARL_S_EXT);
  a[0] = make_float4(vals[0], vals[1], vals[2], vals[3]);
  a[1] = make_float4(vals[4], vals[5], vals[6], vals[7]);
  a[2] = make_float4(vals[8], vals[9], vals[10], vals[11]);
}

// byte occupancy -> bit words (X % 32 == 0, so word w holds voxels 32w..32w+31)
__global__ void k_pack_occ(const uint8_t* __restrict__ bytes, uint32_t* __restrict__ words, size_t nwords) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const uint4* b = reinterpret_cast<const uint4*>(bytes + w * 32);
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const uint4 v = b[q];
    const uint32_t parts[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if ((parts[k] >> (8 * j)) & 0xffu) out |= 1u << (q * 16 + k * 4 + j);
  }
  words[w] = out;
}

__global__ void k_unpack_occ(const uint32_t* __restrict__ words, uint8_t* __restrict__ bytes, size_t nvox) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  bytes[v] = (uint8_t)((words[v >> 5] >> (v & 31)) & 1u);
}

int marl_transpose(marl_ctx* ctx, bool to_soa, float* soa) {
  const size_t n = ctx->nvox;
  const dim3 g((unsigned)((n + 255) / 256));
  if (to_soa)
    hipLaunchKernelGGL(k_aos_to_soa, g, dim3(256), 0, ctx->stream, ctx->staging, soa, n);
  else
    hipLaunchKernelGGL(k_soa_to_aos, g, dim3(256), 0, ctx->stream, soa, ctx->staging, n);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

int marl_pack_occ(marl_ctx* ctx) {
  const size_t nw = ctx->occ_words;
  hipLaunchKernelGGL(k_pack_occ, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, ctx->stream, ctx->occ_bytes,
                     ctx->occ, nw);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

int marl_unpack_occ(marl_ctx* ctx) {
  const size_t n = ctx->nvox;
  hipLaunchKernelGGL(k_unpack_occ, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->occ,
                     ctx->occ_bytes, n);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// Copy probe: a plain float4 streaming copy of known size, used to calibrate
// HBM bandwidth and the FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  // each thread moves 4 float4 spaced one block-width apart (4 loads in flight)
  const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n4) v[k] = a[base + 256 * k];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + 256 * k < n4) b[base + 256 * k] = v[k];
}

int marl_launch_copy_probe(marl_ctx* ctx) {
  const size_t n4 = ctx->nfield / 4;
  hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)((n4 + 1023) / 1024)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float4*>(ctx->field[ctx->cur]),
                     reinterpret_cast<float4*>(ctx->field[ctx->cur ^ 1]), n4);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}
