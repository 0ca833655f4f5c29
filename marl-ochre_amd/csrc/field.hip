// field.hip — field kernels for gfx950: masked 7-point stencil, boundary
// sources, Beer-Lambert light scan, AoS<->SoA transposes, occupancy packing.
//
// Reference semantics:
//   stencil   crates/marl-field/src/field.rs:107-203 (and the WGSL statement
//             crates/marl-gpu/src/shaders/field_diffuse.wgsl:55-95)
//   sources   crates/marl-field/src/field.rs:231-250
//   light     crates/marl-field/src/light.rs:29-59
#include <algorithm>
#include <vector>

#include "marl_internal.h"

// ---------------------------------------------------------------------------
// K1: masked stencil.
//
// A block of 256 threads (4 wave64) owns a 64 (x) x 4*RY (y) tile and marches
// it through a z-chunk of zc planes.  Each thread keeps, for each of its RY
// rows, the 12 species of planes z-1 .. z+PD in registers (a queue refilled
// PD+1 planes ahead); x/y neighbours come from an LDS copy of plane z with a
// one-voxel halo.  Halo rows and columns are prefetched at the same distance
// as the tile's own planes, so the tile and its neighbours (same XCD, same z
// progress) touch the shared cache lines together and the halo hits L2.
// HBM traffic per substep is one read and one write of every species plane
// plus the 1-bit mask, plus one halo plane on each side of a z-chunk.  The
// niche factor is computed in registers from species 7 of the centre voxel.
//
// The per-voxel arithmetic is the reference's, in its order and without
// contraction, so the result is bit-identical to the f32 CPU statement.

// {rows per thread, z-chunk, prefetch distance in planes, species per thread,
//  substeps per pass, waves per block}
const TileShape kTiles[] = {
    // one 4-wave group per block, all 12 species per thread (b128 access)
    {1, 32, 1, 12, 1, 4}, {2, 16, 1, 12, 1, 4}, {2, 32, 1, 12, 1, 4},
    // split blocks: the 4-wave groups of one block share the tile's LDS plane,
    // each group taking 6 (b64 access) or 4 (one quad, b128 access) species
    {2, 16, 1, 6, 1, 8}, {2, 32, 1, 6, 1, 8}, {1, 32, 1, 4, 1, 12}, {2, 16, 1, 4, 1, 12}, {2, 32, 1, 4, 1, 12},
    {1, 64, 1, 4, 1, 12}, {1, 32, 2, 4, 1, 12}, {2, 64, 1, 4, 1, 12}, {1, 128, 1, 4, 1, 12}, {2, 128, 1, 4, 1, 12},
    // 6 species per 4-wave block: the group without species 7 reads it for
    // the niche factor from HBM
    {2, 16, 1, 6, 1, 4},
    // short z-chunks: enough blocks for small grids
    {1, 16, 1, 4, 1, 12}, {1, 8, 1, 4, 1, 12},
    // small grids (field resident in the last-level cache): one quad per
    // 4-wave block, so many small blocks per CU hide the per-plane latency;
    // the quads without species 7 read it for the niche factor
    {1, 8, 1, 4, 1, 4}, {1, 4, 1, 4, 1, 4}, {1, 16, 1, 4, 1, 4}, {1, 8, 2, 4, 1, 4}, {1, 4, 1, 4, 1, 12},
    // temporal blocking (K1b): two substeps per launch
    {1, 16, 1, 12, 2, 8}, {1, 16, 1, 6, 2, 8},
};
const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

// Default tile: the quad-group shapes from the longest z-chunk down, taking
// the first that still gives every CU a block, down to one-quad blocks with
// 4-plane chunks for tiny grids (the autotuner may pick another).
int marl_default_tile(const marl_dims& d) {
  static const int order[] = {11, 10, 8, 7, 5, 14, 15, 17};
  static const int cus = [] {
    int n = 256, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return n;
  }();
  for (int t : order) {
    const TileShape& s = kTiles[t];
    const int zc = s.zc < d.z ? s.zc : d.z;
    const long blocks = (long)((d.x + 63) / 64) * ((d.y + 4 * s.ry - 1) / (4 * s.ry)) * ((d.z + zc - 1) / zc);
    if (blocks * (MARL_S_EXT / (s.sg * (s.nw / 4))) >= cus) return t;
  }
  return 17;
}

// The kernel instantiation of a single-substep tile shape: calls f(TileT<...>)
// for the shape's template parameters (rows, prefetch, species per thread,
// groups per block), or returns hipErrorInvalidValue for a shape without one.
template <int RY_, int PD_, int SG_, int HB_>
struct TileT {
  static constexpr int RY = RY_, PD = PD_, SG = SG_, HB = HB_;
};
template <typename F>
static hipError_t dispatch_tile(const TileShape& t, F&& f) {
  const int hb = t.nw / 4;
#define MARL_TILE(R_, P_, S_, H_) \
  if (t.ry == R_ && t.pd == P_ && t.sg == S_ && hb == H_) return f(TileT<R_, P_, S_, H_>{});
  MARL_TILE(1, 1, 12, 1)
  MARL_TILE(2, 1, 12, 1)
  MARL_TILE(2, 1, 6, 2)
  MARL_TILE(1, 1, 4, 3)
  MARL_TILE(1, 2, 4, 3)
  MARL_TILE(2, 1, 4, 3)
  MARL_TILE(2, 1, 6, 1)
  MARL_TILE(1, 1, 4, 1)
  MARL_TILE(1, 2, 4, 1)
#undef MARL_TILE
  return hipErrorInvalidValue;
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
// non-temporal store (cache policy nt): the substep's output is not re-read
// before the next launch, so it should not displace halo lines in L2
__device__ __forceinline__ void bstore_nt(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 2);
}

// Byte offset of species s within the quad field (voxel and z-plane offsets
// excluded): its quad plane (Qb bytes each) and its lane of the 16-byte quad.
__device__ __forceinline__ uint32_t qoff(int s, uint32_t Qb) { return (uint32_t)(s >> 2) * Qb + (uint32_t)(s & 3) * 4u; }

// A species of a voxel with the boundary sources applied when SRC.
template <bool SRC>
__device__ __forceinline__ float with_sources(float v, const StencilParams& p, int s, int Z, int zz) {
  if (SRC) {
    if (zz == 0) {
      if (s == 1) v = fminf(v + p.src_ox, p.c_max);
      if (s == 3) v = fminf(v + p.src_c, p.c_max);
    }
    if (zz == Z - 1 && s == 2) v = fminf(v + p.src_red, p.c_max);
  }
  return v;
}

// Load species s at plane zz (voff = voxel-in-plane * 16, XYq = bytes per
// z-plane of a quad plane); with SRC the boundary sources of field.rs:231-250
// are applied on the fly (the first substep of a tick reads the sourced field).
template <bool SRC>
__device__ __forceinline__ float ld_field(rsrc_t r, const StencilParams& p, int s, uint32_t Qb, uint32_t XYq,
                                          int Z, uint32_t voff, int zz) {
  return with_sources<SRC>(bload(r, voff, qoff(s, Qb) + (uint32_t)zz * XYq), p, s, Z, zz);
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// A consecutive species s .. s+A-1 of one voxel (s a multiple of A, A = 1, 2
// or 4) with one 4/8/16-byte load: in the quad layout they are adjacent.  The
// wide loads are ordinary global loads from a scalar base (the plane) plus the
// 32-bit voxel offset, not buffer-load builtins: this compiler shrinks a wide
// buffer load to one dword when its elements are re-packed into a vector.
// Offsets stay inside the field (planes are clamped, out-of-grid lanes read
// voxel 0), so no buffer range check is needed.
// The field pointers are cast to the global address space: a pointer that
// reaches the kernel through memory (an ensemble's column table) would
// otherwise be accessed with FLAT instructions, which also count against the
// LDS wait counter, so every LDS wait of the z-loop would wait for the
// prefetches in flight as well.
#define MARL_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const MARL_GLOBAL T* as_global(const void* p) {
  return (const MARL_GLOBAL T*)(p);
}

template <bool SRC, int A>
__device__ __forceinline__ void ld_vec(const float* __restrict__ base, const StencilParams& p, int s, uint32_t Qb,
                                       uint32_t XYq, int Z, uint32_t voff, int zz, float* out) {
  const uint32_t so = qoff(s, Qb) + (uint32_t)zz * XYq;
  const char* a = reinterpret_cast<const char*>(base) + so;
  if constexpr (A == 4) {
    const f32x4 t = *as_global<f32x4>(a + voff);
    out[0] = t.x;
    out[1] = t.y;
    out[2] = t.z;
    out[3] = t.w;
  } else if constexpr (A == 2) {
    const f32x2 t = *as_global<f32x2>(a + voff);
    out[0] = t.x;
    out[1] = t.y;
  } else {
    out[0] = *as_global<float>(a + voff);
  }
#pragma unroll
  for (int k = 0; k < A; ++k) out[k] = with_sources<SRC>(out[k], p, s + k, Z, zz);
}

// Store of species s .. s+A-1 of one voxel with cache policy CP (2: non-temporal).
template <int A, int CP = 2>
__device__ __forceinline__ void st_vec_nt(rsrc_t r, const float* v, uint32_t voff, uint32_t soff) {
  if constexpr (A == 4) {
    const u32x4 t = {__builtin_bit_cast(uint32_t, v[0]), __builtin_bit_cast(uint32_t, v[1]),
                     __builtin_bit_cast(uint32_t, v[2]), __builtin_bit_cast(uint32_t, v[3])};
    __builtin_amdgcn_raw_buffer_store_b128(t, r, (int)voff, (int)soff, CP);
  } else if constexpr (A == 2) {
    const u32x2 t = {__builtin_bit_cast(uint32_t, v[0]), __builtin_bit_cast(uint32_t, v[1])};
    __builtin_amdgcn_raw_buffer_store_b64(t, r, (int)voff, (int)soff, CP);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v[0]), r, (int)voff, (int)soff, CP);
  }
}

// Vector width of a thread's species group: s0 is a multiple of SG.
template <int SG>
constexpr int group_vec() { return SG % 4 == 0 ? 4 : (SG % 2 == 0 ? 2 : 1); }

// LDS of one tile: species planes with a one-voxel halo (one spare row takes
// the writes of lanes that own no halo value, so those writes need no branch)
// and the mask words of the tile's rows.  SB = species held by the block.
template <int RY, int SB>
struct StencilLds {
  float f[SB][4 * RY + 3][66];
  uint32_t m[(4 * RY + 2) * 4 + 256];
};


// One tile (tx, ty, tz) of block species group grp of one substep.  The
// block has HB groups of 4 waves; each takes SG species of the block's HB*SG,
// and all of them share the LDS plane (so a group without species 7 reads the
// niche factor's species 7 from LDS when the block holds it).  With LM = 1
// the tile also runs the light scan on the new field (the chunk must span the
// whole column).
template <int RY, bool SRC, int PD, int SG, int HB = 1, int LM = 0>
__device__ __forceinline__ void stencil_tile(const float* __restrict__ src, float* __restrict__ dst,
                                             const uint32_t* __restrict__ occ, int X, int Y, int Z, int XW, int zc,
                                             int tx, int ty, int tz, int grp, const StencilParams& p,
                                             StencilLds<RY, SG * HB>& sh, float* __restrict__ light_out) {
  constexpr int TY = 4 * RY;
  constexpr int S = SG;              // species handled by this thread
  constexpr int LW = 66;
  auto& lds = sh.f;
  auto& mlds = sh.m;
  // wave group within the block (wave-uniform: kept in a scalar register)
  const int half = HB == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 8));
  const int bs0 = grp * SG * HB;    // first species of this block
  const int s0 = bs0 + half * SG;   // first species of this wave group
  const int ls0 = half * SG;        // its first LDS species slot
  const bool has7 = s0 <= 7 && 7 < s0 + SG;
  const bool blk7 = bs0 <= 7 && 7 < bs0 + SG * HB;  // species 7 is in the block's LDS
  // the fused light scan: the wave group holding species 4 carries each own
  // row's intensity down the column (the chunk is the whole column)
  const bool lit = LM != 0 && s0 <= 4 && 4 < s0 + SG;

  const int tid = threadIdx.x & 255;
  const int lx = tid & 63, wy = tid >> 6;
  const int x0 = tx * 64, y0 = ty * TY, z0 = tz * zc;
  const int z1 = min(z0 + zc, Z);
  const int x = x0 + lx;
  const bool xin = x < X;
  const uint32_t XY = (uint32_t)X * (uint32_t)Y;
  const uint32_t XYq = XY * 16u, Qb = XYq * (uint32_t)Z;  // field bytes per quad z-plane, per quad plane
  const uint32_t XWb = (uint32_t)XW * 4u, MZb = XWb * (uint32_t)Y;
  const rsrc_t rs = make_rsrc(src, Qb * 3u);
  const rsrc_t rd = make_rsrc(dst, Qb * 3u);
  const rsrc_t rm = make_rsrc(occ, MZb * (uint32_t)Z);

  // All loads are unconditional per lane: lanes outside the grid (partial
  // tiles) read element 0 of the plane, a value they never use, so every access
  // stays inside its buffer and there are no branches around loads.
  bool valid[RY];
  int yr[RY];
  uint32_t vo[RY], mo[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    yr[r] = y0 + wy + 4 * r;
    valid[r] = xin && yr[r] < Y;
    vo[r] = valid[r] ? ((uint32_t)yr[r] * (uint32_t)X + (uint32_t)x) * 16u : 0u;
    mo[r] = valid[r] ? (uint32_t)yr[r] * XWb + (uint32_t)(x >> 5) * 4u : 0u;
  }
  // halo slot of this lane: rows y0-1 / y0+TY for waves 0-1, columns x0-1 /
  // x0+64 for the first TY lanes of waves 2-3; other lanes park in the spare row
  int hrow, hcol;
  uint32_t hvo;
  bool hw;
  if (wy < 2) {
    const int yy = (wy == 0) ? y0 - 1 : y0 + TY;
    hw = yy >= 0 && yy < Y && xin;
    hrow = (wy == 0) ? 0 : TY + 1;
    hcol = lx + 1;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)x) * 16u : 0u;
  } else {
    const int xx = (wy == 2) ? x0 - 1 : x0 + 64;
    const int yy = y0 + (lx < TY ? lx : 0);
    hw = lx < TY && xx >= 0 && xx < X && yy < Y;
    hrow = (lx < TY ? lx : 0) + 1;
    hcol = (wy == 2) ? 0 : LW - 1;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)(xx < 0 ? 0 : xx)) * 16u : 0u;
  }
  if (!hw) {
    hrow = TY + 2;
    hcol = lx;
  }
  // mask words of a plane for rows y0-1 .. y0+TY, word columns x0/32-1 .. +2
  const int mrow = tid >> 2, mwc = tid & 3;
  const int myy = y0 - 1 + mrow, mwx = (x0 >> 5) - 1 + mwc;
  const bool mw = tid < (TY + 2) * 4;
  const bool mvalid = mw && myy >= 0 && myy < Y && mwx >= 0 && mwx < XW;
  const uint32_t mvo = mvalid ? (uint32_t)myy * XWb + (uint32_t)mwx * 4u : 0u;
  const int mslot = mw ? tid : (TY + 2) * 4 + tid;

  // Register ring: plane pl of the own rows lives in slot (pl - z0 + 1) mod R,
  // holding planes z-1 .. z+PD+1 at step z.  The z loop is unrolled by R, so
  // every slot index is a compile-time constant: nothing is copied between
  // steps and a prefetch is waited for only where it is first used, one step
  // later, instead of at the end of the step that issued it.
  constexpr int R = PD + 3;
  float q[R][RY][S];
  uint32_t qm[R][RY];
  float h[R][S];    // halo of plane pl (slots of planes z .. z+PD+1)
  uint32_t hm[R];   // this lane's mask word of plane pl
  float q7[R][RY];  // species 7 of plane pl for the niche factor (blocks without it)
  float I[RY];      // light intensity entering plane z (fused light scan)
#pragma unroll
  for (int r = 0; r < RY; ++r) I[r] = p.l_i0;

  constexpr int A = group_vec<SG>();
  auto load_own = [&](int k, int zz) {
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int s = 0; s < S; s += A) ld_vec<SRC, A>(src, p, s0 + s, Qb, XYq, Z, vo[r], zz, &q[k][r][s]);
      qm[k][r] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mo[r], (int)((uint32_t)zz * MZb), 0);
    }
  };
  auto load_halo = [&](int k, int zz) {
#pragma unroll
    for (int s = 0; s < S; s += A) ld_vec<SRC, A>(src, p, s0 + s, Qb, XYq, Z, hvo, zz, &h[k][s]);
    hm[k] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mvo, (int)((uint32_t)zz * MZb), 0);
    if (SG < MARL_S_EXT && !blk7) {
#pragma unroll
      for (int r = 0; r < RY; ++r) q7[k][r] = ld_field<SRC>(rs, p, 7, Qb, XYq, Z, vo[r], zz);
    }
  };
  // prologue: planes z0-1 .. z0+PD (own) and z0 .. z0+PD (halo), clamped into the grid
#pragma unroll
  for (int k = 0; k < PD + 2; ++k) {
    const int zz = z0 - 1 + k;
    load_own(k, zz < 0 ? 0 : (zz < Z ? zz : Z - 1));
    if (k > 0) load_halo(k, zz < Z ? zz : Z - 1);
  }

  for (int zb = z0; zb < z1; zb += R) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int z = zb + u;
      if (z >= z1) break;
      const int km = u % R, kc = (u + 1) % R, kp = (u + 2) % R, kf = (u + PD + 2) % R;
      // prefetch plane z+PD+1 without a branch: past the chunk the own rows
      // re-read plane z1 and the halo plane z1-1 (both already fetched), so
      // the clamp costs no HBM traffic
      const int zf = z + PD + 1;
      const int zo_f = min(min(zf, z1), Z - 1);
      const int zh_f = min(min(zf, z1 - 1), Z - 1);
      load_own(kf, zo_f);
      load_halo(kf, zh_f);
      // plane z into LDS (own values and halo values loaded PD+1 steps ago)
#pragma unroll
      for (int r = 0; r < RY; ++r)
#pragma unroll
        for (int s = 0; s < S; ++s) lds[ls0 + s][wy + 4 * r + 1][lx + 1] = q[kc][r][s];
#pragma unroll
      for (int s = 0; s < S; ++s) lds[ls0 + s][hrow][hcol] = h[kc][s];
      if (half == 0) mlds[mslot] = mvalid ? hm[kc] : 0u;
      __syncthreads();

#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const int row = wy + 4 * r + 1;
        const int y = yr[r];
        auto bitat = [&](int rw, int c) -> uint32_t {
          const int i = c + 32;
          return (mlds[rw * 4 + (i >> 5)] >> (i & 31)) & 1u;
        };
        const uint32_t zo = (uint32_t)z * XYq;
        // neighbour flags with bitwise operators: every mask read is unconditional
        const uint32_t bc = bitat(row, lx), bxm = bitat(row, lx - 1), bxp = bitat(row, lx + 1);
        const uint32_t bym = bitat(row - 1, lx), byp = bitat(row + 1, lx);
        const uint32_t bzm = (qm[km][r] >> (x & 31)) & 1u, bzp = (qm[kp][r] >> (x & 31)) & 1u;
        const bool vxm = (uint32_t)(x > 0) & (bxm ^ 1u);
        const bool vxp = (uint32_t)(x < X - 1) & (bxp ^ 1u);
        const bool vym = (uint32_t)(y > 0) & (bym ^ 1u);
        const bool vyp = (uint32_t)(y < Y - 1) & (byp ^ 1u);
        const bool vzm = (uint32_t)(z > 0) & (bzm ^ 1u);
        const bool vzp = (uint32_t)(z < Z - 1) & (bzp ^ 1u);
        // s0 is a multiple of SG, so species 7 sits at register slot 7 % SG
        const float structural = has7 ? q[kc][r][7 % SG] : (blk7 ? lds[7 - bs0][row][lx + 1] : q7[kc][r]);
        const float niche = 1.0f - p.alpha_eps * structural / (p.k_eps + structural);
        const uint32_t keep = 0u - bc;  // occupied centre: copied unchanged
        float out[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float c = q[kc][r][s];
          const float dd = p.d_voxel[s0 + s] * niche;
          const float lxm = lds[ls0 + s][row][lx];
          const float lxp = lds[ls0 + s][row][lx + 2];
          const float lym = lds[ls0 + s][row - 1][lx + 1];
          const float lyp = lds[ls0 + s][row + 1][lx + 1];
          const float xm = vxm ? lxm : c;
          const float xp = vxp ? lxp : c;
          const float ym = vym ? lym : c;
          const float yp = vyp ? lyp : c;
          const float zm = vzm ? q[km][r][s] : c;
          const float zp = vzp ? q[kp][r][s] : c;
          const float lap = xm + xp + ym + yp + zm + zp - 6.0f * c;
          const float decay = p.lambda_decay[s0 + s] * c;
          const float nc = c + p.dt_sub * (dd * lap - decay);
          // a bit select keeps the compiler from branching around the reads
          out[s] = __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, c) & keep) |
                                                 (__builtin_bit_cast(uint32_t, fmaxf(nc, 0.0f)) & ~keep));
        }
        if (valid[r]) {
          if (LM != 0 && p.view) {
            // into the cell pass's AoS view: species k of voxel v at byte
            // 48 v + 4 k (vo = 16 v, zo = 16 XY z); the block's species groups
            // fill each record's three quads side by side, so the lines are
            // completed in L2 (plain stores, not non-temporal)
            const rsrc_t rv = make_rsrc(p.view, Qb * 3u);
#pragma unroll
            for (int s = 0; s < S; s += A) st_vec_nt<A, 0>(rv, &out[s], vo[r] * 3u, zo * 3u + (uint32_t)(s0 + s) * 4u);
          } else {
#pragma unroll
            for (int s = 0; s < S; s += A) st_vec_nt<A>(rd, &out[s], vo[r], qoff(s0 + s, Qb) + zo);
          }
          // LM 1: light.rs:41-55 on the new species 4 (slot 4 % SG, as for
          // species 7); LM 2: the new species 4 into the light plane, compact,
          // for the separate light scan to read in place
          if (lit) {
            MARL_GLOBAL float* lo = (MARL_GLOBAL float*)light_out;
            if (LM == 1) {
              lo[(size_t)z * XY + (size_t)y * X + x] = I[r];
              if (bc) I[r] *= p.l_e_cell;
              I[r] *= marl_expf(p.l_nchem * out[4 % SG]);
              if (I[r] < p.l_floor) I[r] = 0.0f;
            } else {
              lo[(size_t)z * XY + (size_t)y * X + x] = out[4 % SG];
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

template <int RY, bool SRC, int PD, int SG, int HB, int LM>
__global__ void __launch_bounds__(256 * HB) k_stencil(const float* __restrict__ src, float* __restrict__ dst,
                                                      const uint32_t* __restrict__ occ, int X, int Y, int Z, int XW,
                                                      int zc, int ntx, int nty, int ntz, StencilParams p) {
  __shared__ StencilLds<RY, SG * HB> sh;
  constexpr int NG = MARL_S_EXT / (SG * HB);
  int b = xcd_remap(blockIdx.x, ntx * nty * ntz * NG);
  // species group fastest: the groups of one tile run side by side on one XCD
  const int grp = b % NG;
  b /= NG;
  stencil_tile<RY, SRC, PD, SG, HB, LM>(src, dst, occ, X, Y, Z, XW, zc, b % ntx, (b / ntx) % nty,
                                           b / (ntx * nty), grp, p, sh, p.light);
}

// Ensembles: one launch covers every column of the ensemble.  The logical
// tiles of all columns are numbered back to back (column after column, each
// in the single-column order) and spread over the XCDs by the same remap.
// The grid is either persistent (as many blocks as fit the chip, each walking
// its share with a grid stride) or one block per tile, whichever the
// ensemble's autotuner finds faster; the kernel is the same.
// LM as for k_stencil, into each column's own light plane: 1 (every column's
// z-chunk spans it) runs the light scan, 2 writes the compact species 4.
template <int RY, bool SRC, int PD, int SG, int HB, int LM = 0>
__global__ void __launch_bounds__(256 * HB) k_stencil_ens(const EnsColumn* __restrict__ cols, int ncols, int total,
                                                          int parity, StencilParams p,
                                                          const uint32_t* __restrict__ pieces) {
  __shared__ StencilLds<RY, SG * HB> sh;
  constexpr int NG = MARL_S_EXT / (SG * HB);
  // With a piece table (marl_ens_pieces): block b runs on XCD b % 8 and walks
  // that XCD's pieces [off[x], off[x+1]) with a stride of the XCD's blocks.
  const int x8 = blockIdx.x & 7;
  const bool walk_pieces = pieces != nullptr;
  int w = walk_pieces ? (int)pieces[x8] + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int wend = walk_pieces ? (int)pieces[x8 + 1] : total;
  const int wstep = walk_pieces ? (int)(gridDim.x >> 3) : (int)gridDim.x;
  const uint32_t* table = pieces + kEnsPieceHeader;
  int c = 0;  // column of the previous tile (a block's tiles increase along its walk)
  for (; w < wend; w += wstep) {
    int tx, ty, tz, zc, grp;
    if (walk_pieces) {
      const uint32_t w0 = table[3 * w], w1 = table[3 * w + 1], w2 = table[3 * w + 2];
      c = (int)(w0 & 0xffffu);
      grp = (int)(w0 >> 16);
      tx = (int)(w1 & 0xffffu);
      ty = (int)(w1 >> 16);
      tz = (int)(w2 & 0xffffu);
      zc = (int)(w2 >> 16);
    } else {
      int b = xcd_remap(w, total);
      if (b < cols[c].tile_base) c = 0;
      if (c + 1 < ncols && cols[c + 1].tile_base <= b) {
        int lo = c + 1, hi = ncols - 1;  // binary search: the last column starting at or before b
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (cols[mid].tile_base <= b) lo = mid;
          else hi = mid - 1;
        }
        c = lo;
      }
      const EnsColumn& cb = cols[c];
      b -= cb.tile_base;
      grp = b % NG;
      b /= NG;
      tx = b % cb.ntx;
      ty = (b / cb.ntx) % cb.nty;
      tz = b / (cb.ntx * cb.nty);
      zc = cb.zc;
    }
    const EnsColumn& col = cols[c];
    const int cur = col.cur ^ parity;
    stencil_tile<RY, SRC, PD, SG, HB, LM>(col.f[cur], col.f[cur ^ 1], col.occ, col.X, col.Y, col.Z, col.XW, zc, tx,
                                             ty, tz, grp, p, sh, col.light);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// ---------------------------------------------------------------------------
// K1b: two substeps per pass (temporal blocking).  A block owns an output
// tile of 62 (x) x (TY - 2) (y) voxels and computes the first substep on the
// 64 x TY tile around it (one voxel of overlap on every side).  The
// intermediate field never reaches HBM: each thread keeps its own column's
// intermediate planes z-2, z-1, z in registers and the block keeps plane z-1
// in LDS for the x/y neighbours of the second substep, which runs one plane
// behind the first.  Per voxel both substeps are the single-substep kernel's
// arithmetic in the same order, so the result is bit-identical to two
// launches of K1; HBM traffic per substep is halved (one read and one write
// of the field per two substeps, plus the overlap).
template <int TY>
struct Stencil2Lds {
  float in[MARL_S_EXT][TY + 3][66];   // input plane z with halo (+ spare row)
  float mid[MARL_S_EXT][TY][64];      // intermediate plane z-1
  uint32_t m[2][(TY + 2) * 4 + 64];   // mask words of planes z and z-1 (by parity) + spare slots
};

// NW waves per species group, RY rows per thread (the tile is TY = NW * RY
// rows tall), SG species per thread: 12 / SG groups of NW waves share the
// tile, each handling SG species; niche factors read species 7 from LDS.
template <int RY, bool SRC, int PD, int NW, int SG>
__global__ void __launch_bounds__(64 * NW * (MARL_S_EXT / SG))
    k_stencil2(const float* __restrict__ src, float* __restrict__ dst, const uint32_t* __restrict__ occ, int X,
               int Y, int Z, int XW, int zc, int ntx, int nty, int ntz, StencilParams p) {
  constexpr int S = SG;
  constexpr int TY = NW * RY;
  constexpr int NT = 64 * NW * (MARL_S_EXT / SG);
  constexpr int LW = 66;
  static_assert(TY <= 64 && (TY + 2) * 4 <= NT, "tile rows");
  __shared__ Stencil2Lds<TY> sh;
  auto& lds = sh.in;
  auto& mid = sh.mid;

  int b = xcd_remap(blockIdx.x, ntx * nty * ntz);
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const int tid = threadIdx.x;
  // species group of this wave (wave-uniform: kept in a scalar register)
  const int grp = __builtin_amdgcn_readfirstlane(tid / (64 * NW)), gt = tid % (64 * NW);
  const int s0 = grp * SG;
  const int lx = gt & 63, wy = gt >> 6;
  const int x0 = tx * 62 - 1, y0 = ty * (TY - 2) - 1;  // origin of the first-substep tile
  const int z0 = tz * zc, z1 = min(z0 + zc, Z);          // output planes [z0, z1)
  const int zs = max(z0 - 1, 0), ze = min(z1, Z - 1);    // intermediate planes [zs, ze]
  const int x = x0 + lx;
  const bool xin = x >= 0 && x < X;
  const uint32_t XY = (uint32_t)X * (uint32_t)Y;
  const uint32_t XYq = XY * 16u, Qb = XYq * (uint32_t)Z;  // field bytes per quad z-plane, per quad plane
  const uint32_t XWb = (uint32_t)XW * 4u, MZb = XWb * (uint32_t)Y;
  const rsrc_t rd = make_rsrc(dst, Qb * 3u);
  const rsrc_t rm = make_rsrc(occ, MZb * (uint32_t)Z);

  bool valid[RY], outv[RY];
  int yr[RY];
  uint32_t vo[RY], mo[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int row = wy + NW * r;
    yr[r] = y0 + row;
    valid[r] = xin && yr[r] >= 0 && yr[r] < Y;
    outv[r] = valid[r] && lx >= 1 && lx <= 62 && row >= 1 && row <= TY - 2;
    vo[r] = valid[r] ? ((uint32_t)yr[r] * (uint32_t)X + (uint32_t)x) * 16u : 0u;
    mo[r] = valid[r] ? (uint32_t)yr[r] * XWb + (uint32_t)(x >> 5) * 4u : 0u;
  }
  // halo slot of this lane (as in K1, around the first-substep tile)
  int hrow, hcol;
  uint32_t hvo;
  bool hw;
  if (wy < 2) {
    const int yy = (wy == 0) ? y0 - 1 : y0 + TY;
    hw = yy >= 0 && yy < Y && xin;
    hrow = (wy == 0) ? 0 : TY + 1;
    hcol = lx + 1;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)x) * 16u : 0u;
  } else if (wy >= 4) {
    hw = false;
    hrow = TY + 2;
    hcol = lx;
    hvo = 0u;
  } else {
    const int xx = (wy == 2) ? x0 - 1 : x0 + 64;
    const int yy = y0 + (lx < TY ? lx : 0);
    hw = lx < TY && xx >= 0 && xx < X && yy >= 0 && yy < Y;
    hrow = (lx < TY ? lx : 0) + 1;
    hcol = (wy == 2) ? 0 : LW - 1;
    hvo = hw ? ((uint32_t)yy * (uint32_t)X + (uint32_t)xx) * 16u : 0u;
  }
  if (!hw) {
    hrow = TY + 2;
    hcol = lx;
  }
  // mask words of a plane: rows y0-1 .. y0+TY, word columns wx0-1 .. wx0+2
  const int wx0 = x0 >> 5;  // floor
  const int mrow = gt >> 2, mwc = gt & 3;
  const int myy = y0 - 1 + mrow, mwx = wx0 - 1 + mwc;
  const bool mw = grp == 0 && gt < (TY + 2) * 4;
  const bool mvalid = mw && myy >= 0 && myy < Y && mwx >= 0 && mwx < XW;
  const uint32_t mvo = mvalid ? (uint32_t)myy * XWb + (uint32_t)mwx * 4u : 0u;
  const int mslot = mw ? gt : (TY + 2) * 4 + (tid & 63);  // lanes that write no mask word share spare slots
  // occupancy bit of absolute column xa in mask row rw of buffer mb
  auto bitat = [&](int mb, int rw, int xa) -> uint32_t {
    return (sh.m[mb][rw * 4 + ((xa >> 5) - (wx0 - 1))] >> (xa & 31)) & 1u;
  };

  // Register rings as in K1: input plane pl lives in slot (pl - zs + 1) mod R
  // (planes z-1 .. z+PD+1 at step z), intermediate plane pl in the same slot
  // of p1 (planes z-2, z-1, z); the z loop is unrolled by R so no register
  // holding a prefetch in flight is ever copied.
  constexpr int R = PD + 3;
  static_assert(R >= 3, "intermediate ring");
  float q[R][RY][S];
  uint32_t qm[R][RY];
  float h[R][S];
  uint32_t hm[R];
  float p1[2][RY][S];   // intermediate planes z-2, z-1 of this thread's column (computed, so copied freely)
  uint32_t qmz2[RY];    // mask word of plane z-2 of this thread's column

  constexpr int A = group_vec<SG>();
  auto load_own = [&](int k, int zz) {
#pragma unroll
    for (int r = 0; r < RY; ++r) {
#pragma unroll
      for (int s = 0; s < S; s += A) ld_vec<SRC, A>(src, p, s0 + s, Qb, XYq, Z, vo[r], zz, &q[k][r][s]);
      qm[k][r] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mo[r], (int)((uint32_t)zz * MZb), 0);
    }
  };
  auto load_halo = [&](int k, int zz) {
#pragma unroll
    for (int s = 0; s < S; s += A) ld_vec<SRC, A>(src, p, s0 + s, Qb, XYq, Z, hvo, zz, &h[k][s]);
    hm[k] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)mvo, (int)((uint32_t)zz * MZb), 0);
  };
#pragma unroll
  for (int k = 0; k < PD + 2; ++k) {
    const int zz = zs - 1 + k;
    load_own(k, zz < 0 ? 0 : (zz < Z ? zz : Z - 1));
    if (k > 0) load_halo(k, zz < Z ? zz : Z - 1);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int r = 0; r < RY; ++r)
#pragma unroll
      for (int s = 0; s < S; ++s) p1[k][r][s] = 0.0f;
#pragma unroll
  for (int r = 0; r < RY; ++r) qmz2[r] = 0u;

  for (int zb = zs; zb <= z1; zb += R) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int z = zb + u;
      if (z > z1) break;
      const int km = u % R, kc = (u + 1) % R, kp = (u + 2) % R, kf = (u + PD + 2) % R;
      const bool first = z <= ze;       // this step computes intermediate plane z
      // prefetch input plane z+PD+1 without a branch: own rows are needed up
      // to ze+1 and the halo up to ze; past that the loads repeat a plane
      // already fetched
      const int zf = z + PD + 1;
      load_own(kf, min(min(zf, ze + 1), Z - 1));
      load_halo(kf, min(min(zf, ze), Z - 1));
      if (first) {
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
          for (int s = 0; s < S; ++s) lds[s0 + s][wy + NW * r + 1][lx + 1] = q[kc][r][s];
#pragma unroll
        for (int s = 0; s < S; ++s) lds[s0 + s][hrow][hcol] = h[kc][s];
        sh.m[z & 1][mslot] = mvalid ? hm[kc] : 0u;
      }
      __syncthreads();

      // first substep on plane z (registers only; mid still holds plane z-1)
      float p1n[RY][S];
      if (first) {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
          const int row = wy + NW * r + 1;
          const int y = yr[r];
          const uint32_t bc = bitat(z & 1, row, x), bxm = bitat(z & 1, row, x - 1), bxp = bitat(z & 1, row, x + 1);
          const uint32_t bym = bitat(z & 1, row - 1, x), byp = bitat(z & 1, row + 1, x);
          const uint32_t bzm = (qm[km][r] >> (x & 31)) & 1u, bzp = (qm[kp][r] >> (x & 31)) & 1u;
          const bool vxm = (uint32_t)(x > 0) & (bxm ^ 1u);
          const bool vxp = (uint32_t)(x < X - 1) & (bxp ^ 1u);
          const bool vym = (uint32_t)(y > 0) & (bym ^ 1u);
          const bool vyp = (uint32_t)(y < Y - 1) & (byp ^ 1u);
          const bool vzm = (uint32_t)(z > 0) & (bzm ^ 1u);
          const bool vzp = (uint32_t)(z < Z - 1) & (bzp ^ 1u);
          const float structural = lds[7][row][lx + 1];
          const float niche = 1.0f - p.alpha_eps * structural / (p.k_eps + structural);
          const uint32_t keep = 0u - bc;
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const float c = q[kc][r][s];
            const float dd = p.d_voxel[s0 + s] * niche;
            const float xm = vxm ? lds[s0 + s][row][lx] : c;
            const float xp = vxp ? lds[s0 + s][row][lx + 2] : c;
            const float ym = vym ? lds[s0 + s][row - 1][lx + 1] : c;
            const float yp = vyp ? lds[s0 + s][row + 1][lx + 1] : c;
            const float zm = vzm ? q[km][r][s] : c;
            const float zp = vzp ? q[kp][r][s] : c;
            const float lap = xm + xp + ym + yp + zm + zp - 6.0f * c;
            const float decay = p.lambda_decay[s0 + s] * c;
            const float nc = c + p.dt_sub * (dd * lap - decay);
            p1n[r][s] = __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, c) & keep) |
                                                      (__builtin_bit_cast(uint32_t, fmaxf(nc, 0.0f)) & ~keep));
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
          for (int s = 0; s < S; ++s) p1n[r][s] = 0.0f;
      }

      // second substep on plane zo = z-1 from the intermediate planes
      const int zo = z - 1;
      if (zo >= z0) {
        const int mb = zo & 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
          const int mrw = wy + NW * r;         // row in mid
          const int row = mrw + 1;             // row in the mask LDS
          const int y = yr[r];
          const uint32_t bc = bitat(mb, row, x), bxm = bitat(mb, row, x - 1), bxp = bitat(mb, row, x + 1);
          const uint32_t bym = bitat(mb, row - 1, x), byp = bitat(mb, row + 1, x);
          const uint32_t bzm = (qmz2[r] >> (x & 31)) & 1u, bzp = (qm[kc][r] >> (x & 31)) & 1u;
          const bool vxm = (uint32_t)(x > 0) & (bxm ^ 1u);
          const bool vxp = (uint32_t)(x < X - 1) & (bxp ^ 1u);
          const bool vym = (uint32_t)(y > 0) & (bym ^ 1u);
          const bool vyp = (uint32_t)(y < Y - 1) & (byp ^ 1u);
          const bool vzm = (uint32_t)(zo > 0) & (bzm ^ 1u);
          const bool vzp = (uint32_t)(zo < Z - 1) & (bzp ^ 1u);
          // edge lanes and rows (no output) read clamped slots
          const int cxm = lx > 0 ? lx - 1 : 0, cxp = lx < 63 ? lx + 1 : 63;
          const int rym = mrw > 0 ? mrw - 1 : 0, ryp = mrw < TY - 1 ? mrw + 1 : TY - 1;
          const float structural = mid[7][mrw][lx];
          const float niche = 1.0f - p.alpha_eps * structural / (p.k_eps + structural);
          const uint32_t keep = 0u - bc;
          float out[S];
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const float c = p1[1][r][s];
            const float dd = p.d_voxel[s0 + s] * niche;
            const float xm = vxm ? mid[s0 + s][mrw][cxm] : c;
            const float xp = vxp ? mid[s0 + s][mrw][cxp] : c;
            const float ym = vym ? mid[s0 + s][rym][lx] : c;
            const float yp = vyp ? mid[s0 + s][ryp][lx] : c;
            const float zm = vzm ? p1[0][r][s] : c;
            const float zp = vzp ? p1n[r][s] : c;
            const float lap = xm + xp + ym + yp + zm + zp - 6.0f * c;
            const float decay = p.lambda_decay[s0 + s] * c;
            const float nc = c + p.dt_sub * (dd * lap - decay);
            out[s] = __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, c) & keep) |
                                                   (__builtin_bit_cast(uint32_t, fmaxf(nc, 0.0f)) & ~keep));
          }
          if (outv[r]) {
            const uint32_t zo_b = (uint32_t)zo * XYq;
#pragma unroll
            for (int s = 0; s < S; s += A) st_vec_nt<A>(rd, &out[s], vo[r], qoff(s0 + s, Qb) + zo_b);
          }
        }
      }
      __syncthreads();
      if (first) {
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
          for (int s = 0; s < S; ++s) mid[s0 + s][wy + NW * r][lx] = p1n[r][s];
      }
      // the mask of plane z-1 (fetched steps ago) becomes plane z-2's
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        qmz2[r] = qm[km][r];
#pragma unroll
        for (int s = 0; s < S; ++s) {
          p1[0][r][s] = p1[1][r][s];
          p1[1][r][s] = p1n[r][s];
        }
      }
    }
  }
}

template <int RY, bool SRC, int PD, int SG, int HB>
static hipError_t launch_stencil_t(marl_ctx* ctx, const StencilParams& p, int zc) {
  static_assert(MARL_S_EXT % (SG * HB) == 0, "species split");
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const int ntx = (X + 63) / 64, nty = (Y + 4 * RY - 1) / (4 * RY), ntz = (Z + zc - 1) / zc;
  const int nb = ntx * nty * ntz * (MARL_S_EXT / (SG * HB));
  const float* src = ctx->field[ctx->cur];
  float* dst = ctx->field[ctx->cur ^ 1];
  if (p.fuse_light == 1) {
    if (zc < Z || !p.light) return hipErrorInvalidValue;  // the scan needs whole columns
    hipLaunchKernelGGL((k_stencil<RY, SRC, PD, SG, HB, 1>), dim3(nb), dim3(256 * HB), 0, ctx->stream, src, dst,
                       ctx->occ, X, Y, Z, ctx->xw, zc, ntx, nty, ntz, p);
  } else if (p.fuse_light == 2) {
    if (!p.light) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stencil<RY, SRC, PD, SG, HB, 2>), dim3(nb), dim3(256 * HB), 0, ctx->stream, src, dst,
                       ctx->occ, X, Y, Z, ctx->xw, zc, ntx, nty, ntz, p);
  } else {
    hipLaunchKernelGGL((k_stencil<RY, SRC, PD, SG, HB, 0>), dim3(nb), dim3(256 * HB), 0, ctx->stream, src, dst,
                       ctx->occ, X, Y, Z, ctx->xw, zc, ntx, nty, ntz, p);
  }
  return hipGetLastError();
}

template <int RY, int PD, int SG, int HB = 1>
static hipError_t launch_stencil_s(marl_ctx* ctx, const StencilParams& p, int zc) {
  return p.apply_sources ? launch_stencil_t<RY, true, PD, SG, HB>(ctx, p, zc)
                         : launch_stencil_t<RY, false, PD, SG, HB>(ctx, p, zc);
}

int marl_launch_substep(marl_ctx* ctx, const StencilParams& p, int tile) {
  if (tile < 0 || tile >= kNumTiles || kTiles[tile].tb != 1) tile = marl_default_tile(ctx->d);
  const TileShape t = kTiles[tile];
  const int zc = t.zc < ctx->d.z ? t.zc : ctx->d.z;
  const hipError_t e = dispatch_tile(t, [&](auto T) {
    using TT = decltype(T);
    return launch_stencil_s<TT::RY, TT::PD, TT::SG, TT::HB>(ctx, p, zc);
  });
  MARL_HIP(ctx, e);
  ctx->cur ^= 1;
  return MARL_OK;
}

template <int RY, bool SRC, int PD, int NW, int SG>
static hipError_t launch_stencil2_t(marl_ctx* ctx, const StencilParams& p, int zc) {
  constexpr int TY = NW * RY;
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const int ntx = (X + 61) / 62, nty = (Y + TY - 3) / (TY - 2), ntz = (Z + zc - 1) / zc;
  hipLaunchKernelGGL((k_stencil2<RY, SRC, PD, NW, SG>), dim3(ntx * nty * ntz), dim3(64 * NW * (MARL_S_EXT / SG)), 0,
                     ctx->stream, ctx->field[ctx->cur], ctx->field[ctx->cur ^ 1], ctx->occ, X, Y, Z, ctx->xw, zc, ntx,
                     nty, ntz, p);
  return hipGetLastError();
}

template <int RY, int PD, int NW, int SG>
static hipError_t launch_stencil2_s(marl_ctx* ctx, const StencilParams& p, int zc) {
  return p.apply_sources ? launch_stencil2_t<RY, true, PD, NW, SG>(ctx, p, zc)
                         : launch_stencil2_t<RY, false, PD, NW, SG>(ctx, p, zc);
}

// Two substeps in one launch with a temporal tile.
static int launch_double(marl_ctx* ctx, const StencilParams& p, const TileShape& t) {
  const int zc = t.zc < ctx->d.z ? t.zc : ctx->d.z;
  hipError_t e = hipErrorInvalidValue;
  if (t.sg == 12 && t.ry == 1 && t.pd == 1 && t.nw == 8) e = launch_stencil2_s<1, 1, 8, 12>(ctx, p, zc);
  else if (t.sg == 6 && t.ry == 1 && t.pd == 1 && t.nw == 8) e = launch_stencil2_s<1, 1, 8, 6>(ctx, p, zc);
  MARL_HIP(ctx, e);
  ctx->cur ^= 1;
  return MARL_OK;
}

// The light scan can ride on a single-substep tile whose z-chunk spans the
// column: each thread then walks its (x, y) column from z = 0 down.
static bool tile_fuses_light(int tile, int Z) {
  return tile >= 0 && tile < kNumTiles && kTiles[tile].tb == 1 && kTiles[tile].zc >= Z;
}

// light_mode: what the last substep does for the tick's light scan (0
// nothing, 1 the scan itself, 2 the compact species-4 copy the scan reads).
// With light_mode set, `view` (if not null) takes the last substep's output.
static int substeps_impl(marl_ctx* ctx, const marl_sim_params* sim, int n, bool sources_first, int tile,
                         int* launches, int light_mode, float* view = nullptr) {
  if (tile < 0 || tile >= kNumTiles) tile = marl_default_tile(ctx->d);
  const TileShape t = kTiles[tile];
  int k = 0, nl = 0;
  bool src = sources_first;
  if (t.tb == 2) {
    for (; k + 2 <= n; k += 2, ++nl) {
      if (int rc = launch_double(ctx, marl_stencil_params(sim, src ? 1 : 0), t)) return rc;
      src = false;
    }
  }
  // single substeps (an odd remainder of a temporal tile runs on the default tile)
  const int single = t.tb == 2 ? marl_default_tile(ctx->d) : tile;
  for (; k < n; ++k, ++nl) {
    StencilParams p = marl_stencil_params(sim, src ? 1 : 0);
    if (light_mode && k == n - 1) {
      p.fuse_light = light_mode;
      p.l_i0 = sim->surface_intensity;
      p.l_e_cell = marl_expf(-sim->cell_absorption);
      p.l_nchem = -sim->chemical_absorption;
      p.l_floor = sim->light_floor;
      p.light = ctx->light;
      p.view = view;
    }
    if (int rc = marl_launch_substep(ctx, p, single)) return rc;
    src = false;
  }
  if (launches) *launches = nl;
  return MARL_OK;
}

int marl_launch_substeps(marl_ctx* ctx, const marl_sim_params* sim, int n, bool sources_first, int tile,
                         int* launches) {
  return substeps_impl(ctx, sim, n, sources_first, tile, launches, false);
}

int marl_launch_field_phase(marl_ctx* ctx, const marl_sim_params* sim, bool sources, float* view) {
  const int n = sim->diffusion_substeps;
  ctx->view_fresh = false;
  if (n <= 0) {
    if (sources)
      if (int rc = marl_launch_boundary(ctx, sim)) return rc;
    return marl_launch_light(ctx, sim);
  }
  int tile = ctx->tile;
  if (tile < 0 || tile >= kNumTiles) tile = marl_default_tile(ctx->d);
  // the last substep runs on `tile` itself unless it is a temporal tile (then
  // an odd remainder runs on the default tile, an even count has no single)
  const int last = kTiles[tile].tb == 2 ? (n % 2 ? marl_default_tile(ctx->d) : -1) : tile;
  const int mode = tile_fuses_light(last, ctx->d.z) ? 1 : (last >= 0 ? 2 : 0);
  if (view && !mode) view = nullptr;  // no single last substep to write it
  if (int rc = substeps_impl(ctx, sim, n, sources, ctx->tile, nullptr, mode, view)) return rc;
  ctx->view_fresh = view != nullptr;
  if (mode == 1) return MARL_OK;
  return mode == 2 ? marl_launch_light_compact(ctx, sim) : marl_launch_light(ctx, sim);
}

template <int RY, int PD, int SG, int HB>
static hipError_t launch_stencil_ens(hipStream_t st, const EnsColumn* dcols, int ncols, int total, int parity,
                                     const StencilParams& p, int grid, const uint32_t* pieces) {
  const dim3 g(grid), b(256 * HB);
  if (p.fuse_light == 1) {
    if (p.apply_sources)
      hipLaunchKernelGGL((k_stencil_ens<RY, true, PD, SG, HB, 1>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
    else
      hipLaunchKernelGGL((k_stencil_ens<RY, false, PD, SG, HB, 1>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
  } else if (p.fuse_light == 2) {
    if (p.apply_sources)
      hipLaunchKernelGGL((k_stencil_ens<RY, true, PD, SG, HB, 2>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
    else
      hipLaunchKernelGGL((k_stencil_ens<RY, false, PD, SG, HB, 2>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
  } else if (p.apply_sources) {
    hipLaunchKernelGGL((k_stencil_ens<RY, true, PD, SG, HB>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
  } else {
    hipLaunchKernelGGL((k_stencil_ens<RY, false, PD, SG, HB>), g, b, 0, st, dcols, ncols, total, parity, p, pieces);
  }
  return hipGetLastError();
}

// The persistent ensemble kernel runs the single-substep tiles.
bool marl_tile_ensemble_ok(int tile) {
  return tile >= 0 && tile < kNumTiles && kTiles[tile].tb == 1;
}

// whole_columns: every column's z-chunk is its whole depth (whatever the
// tile's own chunk), which lets the last substep carry the light scan.
void marl_ens_layout(int tile, EnsColumn* cols, int ncols, int* total_tiles, int* total_light_blocks,
                     bool whole_columns, const int* chunks) {
  if (!marl_tile_ensemble_ok(tile)) tile = 7;
  const TileShape t = kTiles[tile];
  const int NG = MARL_S_EXT / (t.sg * (t.nw / 4));
  int base = 0, lbase = 0;
  for (int c = 0; c < ncols; ++c) {
    EnsColumn& e = cols[c];
    const int zc = (chunks && chunks[c] > 0) ? chunks[c] : t.zc;
    e.zc = (whole_columns || zc >= e.Z) ? e.Z : zc;
    e.ntx = (e.X + 63) / 64;
    e.nty = (e.Y + 4 * t.ry - 1) / (4 * t.ry);
    e.ntz = (e.Z + e.zc - 1) / e.zc;
    e.tile_base = base;
    base += e.ntx * e.nty * e.ntz * NG;
    e.lbx = (e.X + 63) / 64;
    e.light_base = lbase;
    lbase += e.lbx * e.Y;
  }
  *total_tiles = base;
  *total_light_blocks = lbase;
}

// Persistent grid: as many blocks as fit the chip at once (occupancy of the
// instantiation x CUs), a multiple of 8 so every XCD gets the same share.
int marl_stencil_ens_grid(int tile, int total_tiles) {
  if (!marl_tile_ensemble_ok(tile)) tile = 7;
  const TileShape t = kTiles[tile];
  int per_cu = 0;
  const hipError_t e = dispatch_tile(t, [&](auto T) {
    using TT = decltype(T);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stencil_ens<TT::RY, false, TT::PD, TT::SG, TT::HB>,
                                                        256 * TT::HB, 0);
  });
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  int grid = per_cu * cus;
  grid = grid < total_tiles ? grid : total_tiles;
  if (grid > 8) grid &= ~7;
  return grid < 1 ? 1 : grid;
}

hipError_t marl_launch_substep_ens(hipStream_t st, const EnsColumn* dcols, int ncols, int total, int parity,
                                   const StencilParams& p, int tile, int grid, const uint32_t* pieces) {
  if (!marl_tile_ensemble_ok(tile)) tile = 7;
  return dispatch_tile(kTiles[tile], [&](auto T) {
    using TT = decltype(T);
    return launch_stencil_ens<TT::RY, TT::PD, TT::SG, TT::HB>(st, dcols, ncols, total, parity, p, grid, pieces);
  });
}

// Balanced pieces for a persistent grid of `grid` blocks (a multiple of 8,
// block b on XCD b % 8).  The stencil tiles of all columns (deepest columns
// first; in each column y, then x, then species group, fastest) are cut into
// 8 runs of equal work (planes), one per XCD, so that the tiles an XCD runs
// side by side are neighbours and share their halo rows in its L2.  In a run,
// the whole-depth tiles that fill complete rounds of the XCD's blocks come
// first, each one piece marching its column's whole depth (no z-chunk halo,
// one pipeline prologue per tile); the tiles of the last, partial round are
// cut into z-chunks so that that round, too, keeps every block busy.
void marl_ens_pieces(int tile, const EnsColumn* cols, int ncols, int grid, std::vector<uint32_t>& out) {
  if (!marl_tile_ensemble_ok(tile)) tile = 7;
  const TileShape t = kTiles[tile];
  const int NG = MARL_S_EXT / (t.sg * (t.nw / 4));
  std::vector<int> ord(ncols);
  for (int c = 0; c < ncols; ++c) ord[c] = c;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cols[a].Z > cols[b].Z; });
  struct Tile {
    int c, tx, ty, grp;
  };
  std::vector<Tile> L;
  int64_t W = 0;
  for (int c : ord) {
    const EnsColumn& e = cols[c];
    const int nty = (e.Y + 4 * t.ry - 1) / (4 * t.ry), ntx = (e.X + 63) / 64;
    for (int ty = 0; ty < nty; ++ty)
      for (int tx = 0; tx < ntx; ++tx)
        for (int g = 0; g < NG; ++g) {
          L.push_back({c, tx, ty, g});
          W += e.Z;
        }
  }
  const int B = grid / 8 > 0 ? grid / 8 : 1;
  out.assign(kEnsPieceHeader, 0u);
  auto emit = [&](const Tile& tl, int tz, int zc) {
    out.push_back((uint32_t)tl.c | (uint32_t)tl.grp << 16);
    out.push_back((uint32_t)tl.tx | (uint32_t)tl.ty << 16);
    out.push_back((uint32_t)tz | (uint32_t)zc << 16);
  };
  size_t pos = 0;
  int64_t acc = 0;
  for (int x = 0; x < 8; ++x) {
    out[x] = (uint32_t)((out.size() - kEnsPieceHeader) / 3);
    const int64_t target = W * (x + 1) / 8;
    const size_t s0 = pos;
    while (pos < L.size() && (x == 7 || acc < target)) acc += cols[L[pos++].c].Z;
    const size_t n = pos - s0, full = n / B * B, rem = n - full;
    for (size_t i = s0; i < s0 + full; ++i) emit(L[i], 0, cols[L[i].c].Z);
    if (rem) {
      const int k = std::max<int>(1, (int)(B / rem));
      for (size_t i = s0 + full; i < pos; ++i) {
        const int Z = cols[L[i].c].Z;
        const int zc = (Z + k - 1) / k;
        for (int tz = 0; tz * zc < Z; ++tz) emit(L[i], tz, zc);
      }
    }
  }
  out[8] = (uint32_t)((out.size() - kEnsPieceHeader) / 3);
}

// ---------------------------------------------------------------------------
// Boundary sources in place (field.rs:231-250): one thread per (x, y).
__global__ void k_boundary(float* __restrict__ f, int X, int Y, int Z, float src_ox, float src_c, float src_red,
                           float c_max) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t XY = (size_t)X * Y, nvox = XY * Z;
  if (i >= (int)XY) return;
  const size_t i1 = fq_index(1, i, nvox), i3 = fq_index(3, i, nvox);
  f[i1] = fminf(f[i1] + src_ox, c_max);
  f[i3] = fminf(f[i3] + src_c, c_max);
  const size_t b = fq_index(2, (size_t)(Z - 1) * XY + i, nvox);
  f[b] = fminf(f[b] + src_red, c_max);
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
// K2: Beer-Lambert light (light.rs:29-59).  One 256-thread block per 64
// (x, y) columns of one row y.  The per-voxel factors do not depend on the
// intensity, so they are computed in parallel: the block's waves evaluate
// expf(-k s4) and the occupancy bit of a chunk of kLightZ planes into LDS,
// and wave 0 then walks the chunk down each column carrying the intensity in
// a register, with the reference's operations in its order per voxel (store
// the incident intensity, times the cell factor if occupied, times the
// chemical factor, floor).  While wave 0 walks chunk k, waves 1-3 fill chunk
// k + 1 into the other LDS buffer.  s4 points at species 4 of voxel 0 in the
// quad field (quad plane 1, lane 0): voxel v's value is s4[4 v].
// ST = 1: s4 is the compact copy the last substep wrote into the light plane
// itself (s4 == light): a chunk's values are read before any of its planes is
// written, and the walk writes only planes already read.
constexpr int kLightZ = 32;
constexpr int kLightThreads = 256;

template <int ST>
__device__ __forceinline__ void light_block(const float* s4, const uint32_t* __restrict__ occ, float* light, int x0,
                                            int y, int X, int Y, int Z, int XW, float I0, float e_cell, float nchem,
                                            float floor_) {
  __shared__ float ex[2][kLightZ][64];
  __shared__ uint8_t ob[2][kLightZ][64];
  // global address space: the ensemble's pointers come from its column table
  const MARL_GLOBAL float* s4g = as_global<float>(s4);
  const MARL_GLOBAL uint32_t* occg = as_global<uint32_t>(occ);
  MARL_GLOBAL float* lg = (MARL_GLOBAL float*)light;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x = x0 + lane;
  const bool xin = x < X;
  const int xc = xin ? x : x0;  // lanes past the row read its first column and write nothing
  const size_t XY = (size_t)X * Y;
  const size_t col = (size_t)y * X + xc;
  const uint32_t bit = 1u << (xc & 31);
  const size_t wcol = (size_t)y * XW + (xc >> 5);
  // planes z0 + j, j = the wave's offset from w0, stepping by the nw waves
  // that fill (4 or 3, so at most kPer planes per wave): every load is
  // issued, from a clamped address, before any value is used
  constexpr int kPer = (kLightZ + 2) / 3;
  auto fill = [&](int buf, int z0, int w0, int nw) {
    float a[kPer];
    uint32_t w[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int z = min(z0 + (wv - w0) + u * nw, Z - 1);
      a[u] = s4g[ST * ((size_t)z * XY + col)];
      w[u] = occg[(size_t)z * Y * XW + wcol];
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int j = (wv - w0) + u * nw;
      if (j < kLightZ && z0 + j < Z) {
        ex[buf][j][lane] = marl_expf(nchem * a[u]);
        ob[buf][j][lane] = (w[u] & bit) ? 1 : 0;
      }
    }
  };
  const int nchunks = (Z + kLightZ - 1) / kLightZ;
  fill(0, 0, 0, kLightThreads / 64);
  __syncthreads();
  float I = I0;
  for (int k = 0; k < nchunks; ++k) {
    const int cur = k & 1;
    if (wv == 0) {
      const int z0 = k * kLightZ, zn = min(kLightZ, Z - z0);
#pragma unroll 8
      for (int j = 0; j < zn; ++j) {
        if (xin) lg[(size_t)(z0 + j) * XY + col] = I;
        if (ob[cur][j][lane]) I *= e_cell;
        I *= ex[cur][j][lane];
        if (I < floor_) I = 0.0f;
      }
    } else if (k + 1 < nchunks) {
      fill(cur ^ 1, (k + 1) * kLightZ, 1, kLightThreads / 64 - 1);
    }
    __syncthreads();
  }
}

template <int ST>
__global__ void __launch_bounds__(kLightThreads) k_light(const float* s4, const uint32_t* __restrict__ occ,
                                                         float* light, int X, int Y, int Z, int XW, float I0,
                                                         float e_cell, float nchem, float floor_) {
  light_block<ST>(s4, occ, light, blockIdx.x * 64, blockIdx.y, X, Y, Z, XW, I0, e_cell, nchem, floor_);
}

// Ensembles: one launch over the (x, y) columns of every column of the ensemble.
template <int ST>
__global__ void __launch_bounds__(kLightThreads) k_light_ens(const EnsColumn* __restrict__ cols, int ncols, float I0,
                                                             float e_cell, float nchem, float floor_) {
  int c = 0;
  const int b = blockIdx.x;
  while (c + 1 < ncols && cols[c + 1].light_base <= b) ++c;
  const EnsColumn& e = cols[c];
  const int local = b - e.light_base;
  const size_t nvox = (size_t)e.X * e.Y * e.Z;
  light_block<ST>(ST == 1 ? e.light : e.f[e.cur] + 4 * nvox, e.occ, e.light, (local % e.lbx) * 64, local / e.lbx,
                  e.X, e.Y, e.Z, e.XW, I0, e_cell, nchem, floor_);
}

hipError_t marl_launch_light_ens(hipStream_t st, const EnsColumn* dcols, int ncols, int total_blocks,
                                 const marl_sim_params* sim, bool compact) {
  const float e_cell = marl_expf(-sim->cell_absorption);
  if (compact)
    hipLaunchKernelGGL(k_light_ens<1>, dim3(total_blocks), dim3(kLightThreads), 0, st, dcols, ncols, sim->surface_intensity,
                       e_cell, -sim->chemical_absorption, sim->light_floor);
  else
    hipLaunchKernelGGL(k_light_ens<4>, dim3(total_blocks), dim3(kLightThreads), 0, st, dcols, ncols, sim->surface_intensity,
                       e_cell, -sim->chemical_absorption, sim->light_floor);
  return hipGetLastError();
}

int marl_launch_light(marl_ctx* ctx, const marl_sim_params* sim) {
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const float* s4 = ctx->field[ctx->cur] + (size_t)4 * ctx->nvox;
  const float e_cell = marl_expf(-sim->cell_absorption);
  hipLaunchKernelGGL(k_light<4>, dim3((X + 63) / 64, Y), dim3(kLightThreads), 0, ctx->stream, s4, ctx->occ, ctx->light, X, Y,
                     Z, ctx->xw, sim->surface_intensity, e_cell, -sim->chemical_absorption, sim->light_floor);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// The light scan on the compact species 4 the tick's last substep left in the
// light plane (read and overwritten in place).
int marl_launch_light_compact(marl_ctx* ctx, const marl_sim_params* sim) {
  const int X = ctx->d.x, Y = ctx->d.y, Z = ctx->d.z;
  const float e_cell = marl_expf(-sim->cell_absorption);
  hipLaunchKernelGGL(k_light<1>, dim3((X + 63) / 64, Y), dim3(kLightThreads), 0, ctx->stream, ctx->light, ctx->occ, ctx->light,
                     X, Y, Z, ctx->xw, sim->surface_intensity, e_cell, -sim->chemical_absorption, sim->light_floor);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

// ---------------------------------------------------------------------------
// Layout kernels.  AoS (the boundary's field.rs layout, 12 floats per voxel)
// and the resident quads hold the same 16-byte groups, so the conversion is a
// float4 permutation; species planes (SoA) go through per-float moves.
// One float4 per thread, in AoS order: the AoS side is a linear stream and the
// quad side three interleaved runs (one per plane) of consecutive voxels.
// 3 * nvox < 2^32 (the field fits 2 GiB).
__global__ void k_aos_to_quad(const float4* __restrict__ aos, float4* __restrict__ q, uint32_t nvox) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3u * nvox) return;
  const uint32_t v = i / 3u, g = i - 3u * v;
  q[(size_t)g * nvox + v] = aos[i];
}

__global__ void k_quad_to_aos(const float4* __restrict__ q, float4* __restrict__ aos, uint32_t nvox) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3u * nvox) return;
  const uint32_t v = i / 3u, g = i - 3u * v;
  aos[i] = q[(size_t)g * nvox + v];
}

__global__ void k_soa_to_quad(const float* __restrict__ soa, float4* __restrict__ q, size_t nvox) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
#pragma unroll
  for (int g = 0; g < 3; ++g)
    q[(size_t)g * nvox + v] = make_float4(soa[(size_t)(4 * g) * nvox + v], soa[(size_t)(4 * g + 1) * nvox + v],
                                          soa[(size_t)(4 * g + 2) * nvox + v], soa[(size_t)(4 * g + 3) * nvox + v]);
}

__global__ void k_quad_to_soa(const float4* __restrict__ q, float* __restrict__ soa, size_t nvox) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const float4 a = q[(size_t)g * nvox + v];
    soa[(size_t)(4 * g) * nvox + v] = a.x;
    soa[(size_t)(4 * g + 1) * nvox + v] = a.y;
    soa[(size_t)(4 * g + 2) * nvox + v] = a.z;
    soa[(size_t)(4 * g + 3) * nvox + v] = a.w;
  }
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

int marl_transpose(marl_ctx* ctx, bool to_quads, float* quads) {
  const uint32_t n = (uint32_t)ctx->nvox;
  const dim3 g((3u * n + 255u) / 256u);
  float4* q = reinterpret_cast<float4*>(quads);
  float4* a = reinterpret_cast<float4*>(ctx->staging);
  if (to_quads)
    hipLaunchKernelGGL(k_aos_to_quad, g, dim3(256), 0, ctx->stream, a, q, n);
  else
    hipLaunchKernelGGL(k_quad_to_aos, g, dim3(256), 0, ctx->stream, q, a, n);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}

int marl_planes(marl_ctx* ctx, bool to_quads, float* quads) {
  const size_t n = ctx->nvox;
  const dim3 g((unsigned)((n + 255) / 256));
  float4* q = reinterpret_cast<float4*>(quads);
  if (to_quads)
    hipLaunchKernelGGL(k_soa_to_quad, g, dim3(256), 0, ctx->stream, ctx->staging, q, n);
  else
    hipLaunchKernelGGL(k_quad_to_soa, g, dim3(256), 0, ctx->stream, q, ctx->staging, n);
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
// HBM bandwidth and the FETCH_SIZE / WRITE_SIZE counters.  One float4 per
// thread: of the copy shapes measured (grid strides, 4 or 8 float4 per
// thread, nontemporal loads and stores) this one moves the most, ~6.2 TB/s.
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

// Gather probe: the random-access ceiling for scattered-record kernels (the
// cell pass).  Every 4 consecutive lanes read one 64-byte sector of the
// resident field at a hashed position, four sectors per thread, and each
// thread writes one float of the spare buffer so the reads are kept.
// Positions are a hash modulo the sector count, so every read is in bounds.
__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

constexpr int kGatherPerThread = 4;

__global__ void __launch_bounds__(256) k_gather_probe(const float4* __restrict__ a, uint32_t nsectors,
                                                      float* __restrict__ out, uint32_t salt) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  float acc = 0.0f;
#pragma unroll
  for (int r = 0; r < kGatherPerThread; ++r) {
    const uint32_t s = probe_hash(((t >> 2) * kGatherPerThread + r) ^ salt) % nsectors;
    const float4 v = a[(size_t)s * 4 + (t & 3)];
    acc += v.x + v.y + v.z + v.w;
  }
  out[t] = acc;
}

// src: nsectors 64-byte sectors; out: one float per thread (threads a
// multiple of 256).
int marl_launch_gather_probe(marl_ctx* ctx, const void* src, uint32_t nsectors, float* out, size_t threads,
                             uint32_t salt, double* sector_bytes) {
  hipLaunchKernelGGL(k_gather_probe, dim3((unsigned)(threads / 256)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float4*>(src), nsectors, out, salt);
  MARL_HIP(ctx, hipGetLastError());
  if (sector_bytes) *sector_bytes = (double)(threads / 4) * kGatherPerThread * 64.0;
  return MARL_OK;
}

int marl_launch_copy_probe(marl_ctx* ctx) {
  const size_t n4 = ctx->nfield / 4;
  hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float4*>(ctx->field[ctx->cur]),
                     reinterpret_cast<float4*>(ctx->field[ctx->cur ^ 1]), n4);
  MARL_HIP(ctx, hipGetLastError());
  return MARL_OK;
}
