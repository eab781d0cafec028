// Acoustic p-u-w staggered-grid propagator for gfx950: fused forward step,
// fused discrete-adjoint step, receiver gather / adjoint-source injection,
// and the C ABI around them.
//
// One forward step (reference: ADFWI/propagator/acoustic_kernels.py:113-174):
//   1. p <- (1-K1) p - A1 (c1 (u_j - u_{j-1} + w_i - w_{i-1}) + c2 (u_{j+1} - u_{j-2} + w_{i+1} - w_{i-2}))
//      on rows [fs+1, Z-3] x cols [2, X-3]
//   2. p[zs, xs] += dt * wavelet[t]
//   3. free surface: p[fs-1, :] = -p[fs+1, :]
//   4. u <- (1-K2) u - A2 (c1 (p_{j+1} - p_j) + c2 (p_{j+2} - p_{j-1}))  on [fs, Z-2] x [1, X-3]
//   5. w <- (1-K3) w - A2 (c1 (p_{i+1} - p_i) + c2 (p_{i+2} - p_{i-1}))  on [fs, Z-3] x [1, X-2]
//   6. free surface: w[fs-1, :] = w[fs, :]
//   7. records of p, u, w; 8. energy maps.
// Every expression keeps the reference's fp32 association (the library is
// compiled with -ffp-contract=off), so a step is bit-identical to the
// reference's CPU op sequence given identical coefficient fields.
//
// Work decomposition ("strips"): one wavefront owns 64 consecutive columns
// x0-3 .. x0+60 of a TZ-row strip and writes the 58 interior columns
// x0 .. x0+57; the 3 + 3 halo lanes recompute the new pressure their
// neighbours need.  Each lane marches down its column keeping the vertical
// stencil windows in registers; horizontal neighbours come from cross-lane
// shuffles, so a point costs one coalesced load/store per field.  A wave loops
// over the shots of its shot group, accumulating the energy maps (forward) or
// the coefficient gradients (adjoint) in registers before one read-modify-
// write of the group's accumulator plane.  State is ping-ponged between two
// buffer sets, so halo reads never race with neighbouring strips' writes.
#include "common.hpp"

namespace {

constexpr int WAVE = 64;
constexpr int HALO_L = 3;                     // lanes left of the first output column
constexpr int SX = WAVE - 6;                  // output columns per strip (58)
constexpr int TZ_MAX = 64;                    // output rows per strip: runtime, 8..64
constexpr int WPB = 4;                        // wavefronts (strips) per workgroup
constexpr int NT = WAVE * WPB;

constexpr float C1 = 9.0f / 8.0f;
constexpr float C2 = static_cast<float>(-1.0 / 24.0);

struct AcDev {
  int S, Z, X, pitch, fs, fsurf, nt, nrec;
  int zbase, nsx, nsz, nstrips, spb, unroll, tz;
  long long plane;  // Z * pitch
  long long field;  // S * Z * pitch
  const float *coef;
  const float *src;
  const int *src_z, *src_x, *rcv_z, *rcv_x;
  // forward only: device-side time base added to every step index (NULL: 0), so that one
  // captured launch sequence serves every segment of a run; first step of the u energy map
  const int *tbase;
  int eu_from;
  // per-shot coefficient set (NULL: one set for all shots); all shots of a group share one
  const int *coef_set;
};

// the five coefficient planes of the shot group that starts at shot s0
__device__ __forceinline__ const float *coef_of(const AcDev &d, int s0) {
  return d.coef_set ? d.coef + (long long)d.coef_set[s0] * 5 * d.plane : d.coef;
}

// the time base of a forward launch (one scalar load; 0 without a device time base)
__device__ __forceinline__ int time_base(const AcDev &d) { return d.tbase ? *d.tbase : 0; }

__device__ __forceinline__ float shfl(float v, int lane) { return __shfl(v, lane, WAVE); }

// Raw buffer access: a scalar resource descriptor + a wave-uniform byte offset (SGPR) + the
// lane's column byte offset (VGPR), so no per-lane 64-bit address arithmetic is needed.
// Every access stays inside its buffer by construction: loads use clamped rows and the clamped
// column (vc), stores use the lane's own column (vs) only on lanes that own an output point.
// The hardware range check is relied upon for absent buffers (num_records 0) and, in the
// branch-free two-step variants, for accesses that a lane must not make: their lane offset
// DROP points past the buffer, so the access is dropped, moves no data and reads 0.  That
// needs every descriptor under 2 GiB (DROP plus the scalar offset then stays below 2^32 and
// at or above num_records); the host picks those variants only then (pair_drop_ok).  Stores
// issued on every lane that way keep the loop free of branches around them, whose skipped
// path would make the loop's wait for its prefetched loads also wait for the stores.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int RSRC_DW3 = 0x00020000;  // raw 32-bit data, gfx9 family
__device__ __forceinline__ rsrc_t rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_DW3);
}
// the b32 builtins move raw 32-bit words: bit-cast, never value-convert
__device__ __forceinline__ float bld(rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(float v, rsrc_t r, unsigned voff, unsigned soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
constexpr unsigned DROP = 0x80000000u;

// strip index -> (z0, x0).  Workgroups are dispatched round-robin to the 8 XCDs
// (each with its own L2); this bijection gives every XCD a contiguous run of
// strips so that halo rows and coefficient lines are shared inside one L2.
__device__ __forceinline__ bool strip_of(const AcDev &d, int &z0, int &x0) {
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int xcd = b & 7, per = n >> 3, rem = n & 7;
  const int logical = xcd * per + min(xcd, rem) + (b >> 3);
  // the wave index is uniform across the wavefront: keep it (and all row offsets) scalar
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int strip = logical * WPB + wv;
  if (strip >= d.nstrips) return false;
  const int sz = strip / d.nsx;
  z0 = d.zbase + sz * d.tz;
  x0 = (strip - sz * d.nsx) * SX;
  return true;
}

// ---------------------------------------------------------------------------
// forward step
// ---------------------------------------------------------------------------
template <int SPB, int UNR>
__global__ __launch_bounds__(NT) void ac_fwd_step(
    AcDev d, const float *__restrict__ in, float *__restrict__ out, int t,
    float *__restrict__ energy, int acc_u, float *__restrict__ img_d, float *__restrict__ img_p) {
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;
  t += time_base(d);
  if (t < 0 || t >= d.nt) return;      // a device time base outside the run: nothing to do
  const bool want_u = acc_u && t >= d.eu_from;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - HALO_L + lane;                 // this lane's column
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);   // clamped load column (bytes)
  const unsigned vs = 4u * (unsigned)max(c, 0);                 // store column (output lanes only)
  const bool out_col = lane >= HALO_L && lane < HALO_L + SX && c < d.X;
  const bool cp = c >= 2 && c <= d.X - 3;           // column part of the windows
  const bool cu = c >= 1 && c <= d.X - 3;
  const bool cw = c >= 1 && c <= d.X - 2;
  const bool u_ok = c < d.X - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch;
  const unsigned planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  const int s0 = blockIdx.y * SPB;
  const int ns = min(SPB, d.S - s0);
  const int zmax = d.Z - 1;
  const int fs = d.fs;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rP = rsrc(in, fieldb), rU = rsrc(in + d.field, fieldb), rW = rsrc(in + 2 * d.field, fieldb);
  const rsrc_t oP = rsrc(out, fieldb), oU = rsrc(out + d.field, fieldb), oW = rsrc(out + 2 * d.field, fieldb);
  const rsrc_t rC = rsrc(coef_of(d, s0), 5u * planeb);
  const rsrc_t rI = rsrc(img_d, img_d ? fieldb : 0u), rIp = rsrc(img_p, img_p ? fieldb : 0u);
  const rsrc_t rE = rsrc(energy, energy ? 2u * planeb * gridDim.y : 0u);
  const unsigned eb = 2u * planeb * blockIdx.y;
  unsigned so[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) so[j] = (unsigned)(s0 + min(j, ns - 1)) * planeb;

  // per-shot windows: pw* = new p rows k-3..k; ww* = old w rows k-2..k+1; uw* = old u rows k-2..k
  float pw0[SPB], pw1[SPB], pw2[SPB], pw3[SPB];
  float ww0[SPB], ww1[SPB], ww2[SPB];
  float uw0[SPB], uw1[SPB];
  // values of the row being processed, loaded one iteration ahead: W(k+1), U(k), P(k)
  float cW[SPB], cU[SPB], cP[SPB];
  float sv[SPB];
  int zs[SPB], xs[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) {
    pw0[j] = pw1[j] = pw2[j] = pw3[j] = 0.0f;
    uw0[j] = uw1[j] = 0.0f;
    const int s = s0 + min(j, ns - 1);
    zs[j] = j < ns ? d.src_z[s] : -1000;
    xs[j] = d.src_x[s];
    sv[j] = d.src[(long long)s * d.nt + t];
    ww0[j] = 0.0f;
    ww1[j] = bld(rW, vc, so[j] + rb(z0 - 3));
    ww2[j] = bld(rW, vc, so[j] + rb(z0 - 2));
    cW[j] = bld(rW, vc, so[j] + rb(z0 - 1));   // becomes w(k+1) once shifted for k = z0-1 ... see below
  }
  // first row k = z0-1 needs w rows z0-3..z0: shift so that ww* hold rows k-2..k and cW holds k+1
#pragma unroll
  for (int j = 0; j < SPB; ++j) {
    ww0[j] = ww1[j]; ww1[j] = ww2[j]; ww2[j] = cW[j];
    cW[j] = bld(rW, vc, so[j] + rb(z0));
    cU[j] = bld(rU, vc, so[j] + rb(z0 - 1));
    cP[j] = bld(rP, vc, so[j] + rb(z0 - 1));
  }
  float ck1 = bld(rC, vc, rb(z0 - 1)), ca1 = bld(rC, vc, planeb + rb(z0 - 1));
  float ck2 = bld(rC, vc, 2u * planeb + rb(z0 - 3)), ca2 = bld(rC, vc, 3u * planeb + rb(z0 - 3));
  float ck3 = bld(rC, vc, 4u * planeb + rb(z0 - 3));
  // energy rows are read-modify-written once per row: their old values are loaded one
  // iteration ahead too, so that no load is waited for in the iteration that issues it
  float ce0 = energy ? bld(rE, vc, eb + rb(z0 - 3)) : 0.0f;
  float ce1 = (energy && want_u) ? bld(rE, vc, eb + planeb + rb(z0 - 3)) : 0.0f;

#pragma unroll UNR
  for (int kk = 0; kk < d.tz + 3; ++kk) {
    const int k = z0 - 1 + kk;
    const unsigned rk = rb(k);
    const int r = k - 2;
    const unsigned rr = rb(r);
    // issue the next row's loads (all shots) before this row's stores
    float nW[SPB], nU[SPB], nP[SPB];
    const unsigned rn = rb(k + 1), rn1 = rb(k + 2), rrn = rb(r + 1);
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      nW[j] = bld(rW, vc, so[j] + rn1);
      nU[j] = bld(rU, vc, so[j] + rn);
      nP[j] = bld(rP, vc, so[j] + rn);
    }
    const float nk1 = bld(rC, vc, rn), na1 = bld(rC, vc, planeb + rn);
    const float nk2 = bld(rC, vc, 2u * planeb + rrn), na2 = bld(rC, vc, 3u * planeb + rrn);
    const float nk3 = bld(rC, vc, 4u * planeb + rrn);
    const float ne0 = energy ? bld(rE, vc, eb + rrn) : 0.0f;
    const float ne1 = (energy && want_u) ? bld(rE, vc, eb + planeb + rrn) : 0.0f;

    const bool rowp = cp && k >= fs + 1 && k <= d.Z - 3;
    const bool rowu = cu && r >= fs && r <= d.Z - 2;
    const bool roww = cw && r >= fs && r <= d.Z - 3;
    const bool outr = kk >= 3;
    const bool own = out_col && kk >= 1 && kk <= d.tz && k <= zmax;   // row k belongs to this strip
    const bool wout = out_col && outr && r <= zmax;
    float ep = 0.0f, eu = 0.0f;
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      if (j < ns) {
        // windows: w rows k-2..k+1 = ww0, ww1, ww2, cW; u rows k-2..k = uw0, uw1, cU
        const float w3 = cW[j];
        float pn = cP[j];
        // stage 1: new pressure at row k
        const float u0 = cU[j];
        const float um2 = shfl(u0, lane - 2), um1 = shfl(u0, lane - 1), up1 = shfl(u0, lane + 1);
        float dd = 0.0f;
        if (rowp) {
          dd = C1 * (((u0 - um1) + ww2[j]) - ww1[j]) + C2 * (((up1 - um2) + w3) - ww0[j]);
          pn = ck1 * pn - ca1 * dd;
        }
        // stage 2: source; stage 3: free-surface mirror of row fs-1 (two rows back)
        if (k == zs[j] && c == xs[j]) pn = pn + sv[j];
        pw0[j] = pw1[j]; pw1[j] = pw2[j]; pw2[j] = pw3[j]; pw3[j] = pn;
        if (d.fsurf && k == fs + 1) pw1[j] = -pn;
        if (img_d && own) bst(dd, rI, vs, so[j] + rk);
        if (outr) {
          // stages 4-6 at output row r = k-2 (u(r) = uw0, w(r) = ww0)
          const float pr = pw1[j];
          const float pl1 = shfl(pr, lane - 1), pr1 = shfl(pr, lane + 1), pr2 = shfl(pr, lane + 2);
          float un = uw0[j], wn = ww0[j];
          if (rowu) un = ck2 * un - ca2 * (C1 * (pr1 - pr) + C2 * (pr2 - pl1));
          if (roww) wn = ck3 * wn - ca2 * (C1 * (pw2[j] - pr) + C2 * (pw3[j] - pw0[j]));
          if (wout) {
            bst(pr, oP, vs, so[j] + rr);
            if (img_p) bst(pr, rIp, vs, so[j] + rr);
            if (u_ok) bst(un, oU, vs, so[j] + rr); else un = 0.0f;
            // stage 6: row fs-1 takes the new row fs, written once that row exists
            if (r < d.Z - 1 && !(d.fsurf && r == fs - 1)) bst(wn, oW, vs, so[j] + rr);
            if (d.fsurf && r == fs) bst(wn, oW, vs, so[j] + rr - rowb);
            ep += pr * pr;
            eu += un * un;
          }
        }
        // advance the old-state windows by one row
        ww0[j] = ww1[j]; ww1[j] = ww2[j]; ww2[j] = w3;
        uw0[j] = uw1[j]; uw1[j] = u0;
      }
      cW[j] = nW[j]; cU[j] = nU[j]; cP[j] = nP[j];
    }
    ck1 = nk1; ca1 = na1; ck2 = nk2; ca2 = na2; ck3 = nk3;
    if (energy && wout) {
      bst(ce0 + ep, rE, vs, eb + rr);
      if (want_u) bst(ce1 + eu, rE, vs, eb + planeb + rr);
    }
    ce0 = ne0; ce1 = ne1;
  }
}

// ---------------------------------------------------------------------------
// two forward steps in one launch (temporal blocking)
// ---------------------------------------------------------------------------
// Step A (t) runs exactly as ac_fwd_step: at iteration k it forms p'(k) and emits the finished
// row r = k-2 of the intermediate state (p', u', w').  Step B (t+1) consumes those rows from
// registers, three rows behind: p''(q) at q = k-3 and u'', w'' at q' = k-5, and only B's rows are
// stored.  The intermediate state never touches HBM; its receiver samples are written when
// their row is emitted (through a receiver-index map), its energy terms wait three rows in
// registers so that one read-modify-write adds both steps.  Every lane needs 6 columns of halo
// on each side (3 per step), so a wavefront writes 52 of its 64 columns.
constexpr int HALO2 = 6;
constexpr int SX2 = WAVE - 2 * HALO2;   // 52

// feature sets of the two-step forward kernel: records and energy maps, stencil images,
// pressure images
constexpr int PF_REC = 1, PF_IMGD = 2, PF_IMGP = 4;
// PF_BRANCH: stores in branches and receiver slots read in their own iteration, with no
// dropped accesses (for fields whose two-field descriptors would reach 2 GiB)
constexpr int PF_BRANCH = 8;

struct PairIO {
  const float *in;
  float *out;
  int t;
  float *energy;
  float *img_d0, *img_d1, *img_p0, *img_p1;   // *_1 = *_0 + one field (consecutive steps)
  const int *rmap;                      // [Z] receivers per row, then [Z][pitch] receiver index or -1
  float *rec_p, *rec_u, *rec_w;         // records; step t is written here
};

__device__ __forceinline__ bool strip_of2(const AcDev &d, int &z0, int &x0) {
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int xcd = b & 7, per = n >> 3, rem = n & 7;
  const int logical = xcd * per + min(xcd, rem) + (b >> 3);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int strip = logical * WPB + wv;
  if (strip >= d.nstrips) return false;
  const int sz = strip / d.nsx;
  z0 = d.zbase + sz * d.tz;
  x0 = (strip - sz * d.nsx) * SX2;
  return true;
}

template <int SPB, int F>
__global__ __launch_bounds__(NT) void ac_fwd_pair(AcDev d, PairIO io) {
  int z0, x0;
  if (!strip_of2(d, z0, x0)) return;
  // features compiled out are null here, so their descriptors and branches fold away and
  // the scalar registers they would hold stay free
  constexpr bool FR = (F & PF_REC) != 0, FD = (F & PF_IMGD) != 0, FP = (F & PF_IMGP) != 0;
  float *const energy = FR ? io.energy : nullptr;
  float *const rec_p = FR ? io.rec_p : nullptr;
  float *const img_d0 = FD ? io.img_d0 : nullptr, *const img_p0 = FP ? io.img_p0 : nullptr;
  const bool img_d1 = img_d0 != nullptr, img_p1 = img_p0 != nullptr;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - HALO2 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= HALO2 && lane < HALO2 + SX2 && c < d.X;
  const bool cp = c >= 2 && c <= d.X - 3;
  const bool cu = c >= 1 && c <= d.X - 3;
  const bool cw = c >= 1 && c <= d.X - 2;
  const bool u_ok = c < d.X - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch;
  const unsigned planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  const int s0 = blockIdx.y * SPB;
  const int ns = min(SPB, d.S - s0);
  const int zmax = d.Z - 1;
  const int fs = d.fs;
  const int end_row = min(z0 + d.tz, d.Z);
  const int t = io.t + time_base(d);
  if (t < 0 || t + 1 >= d.nt) return;  // a device time base outside the run: nothing to do
  const bool want_u0 = t >= d.eu_from, want_u1 = t + 1 >= d.eu_from;   // u^2 enters the energy map
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  // one descriptor spans two consecutive fields (p and u; the images of steps t and t+1): the
  // second field is reached through the lane offset (vc1 / vs1), which keeps scalar registers
  // free.  A field is under 2 GiB, so two fit the 32-bit offsets.
  const unsigned vc1 = vc + fieldb, vs1 = vs + fieldb;
  const rsrc_t rPU = rsrc(io.in, 2u * fieldb), rW = rsrc(io.in + 2 * d.field, fieldb);
  const rsrc_t oPU = rsrc(io.out, 2u * fieldb), oW = rsrc(io.out + 2 * d.field, fieldb);
  const rsrc_t rC = rsrc(coef_of(d, s0), 5u * planeb);
  const rsrc_t rI = rsrc(img_d0, img_d0 ? 2u * fieldb : 0u), rJ = rsrc(img_p0, img_p0 ? 2u * fieldb : 0u);
  const rsrc_t rE = rsrc(energy, energy ? 2u * planeb * gridDim.y : 0u);
  const unsigned eb = 2u * planeb * blockIdx.y;
  const bool recs = rec_p != nullptr && d.nrec > 0;
  // rows z0 .. z0+127 that hold receivers (a strip's records are of rows z0 .. end_row-1)
  const adfwi::RowMask rrows = adfwi::receiver_rows(recs ? io.rmap : nullptr, d.Z, z0, lane);
  // BF: stores issued on every lane (dropped where a lane owns no point), in the recompute
  // variants (images, no records).  The variants with records keep their stores in branches
  // and read a receiver slot in its own iteration: with records and images the registers would
  // cost a wave per SIMD, and with records only it measured 1.7 % slower (334 against 328 us;
  // the recompute variant 349 against 367 us, one box)
  constexpr bool BF = (F & PF_BRANCH) == 0 && !FR;
  // a store that only some lanes make
  auto put = [&](bool own, float v, rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (BF) bst(v, r, own ? voff : DROP, soff);
    else if (own) bst(v, r, voff, soff);
  };
  unsigned so[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) so[j] = (unsigned)(s0 + min(j, ns - 1)) * planeb;

  // step A windows (as ac_fwd_step): pw* new p rows k-3..k; ww* old w rows k-2..k, cW row k+1;
  // uw* old u rows k-2..k-1, cU row k; cP old p row k
  float pw0[SPB], pw1[SPB], pw2[SPB], pw3[SPB];
  float ww0[SPB], ww1[SPB], ww2[SPB];
  float uw0[SPB], uw1[SPB];
  float cW[SPB], cU[SPB], cP[SPB];
  // step B windows over A's emitted rows (r = k-2): bw* w' rows r-3..r, bu* u' rows r-3..r,
  // bpo p' row r-1; qw* new p'' rows q-3..q
  float bw0[SPB], bw1[SPB], bw2[SPB], bw3[SPB];
  float bu0[SPB], bu1[SPB], bu2[SPB], bu3[SPB];
  float bpo[SPB];
  float qw0[SPB], qw1[SPB], qw2[SPB], qw3[SPB];
  // sources: the row of each shot (uniform) and a per-lane bit field of the shots whose source
  // column is this lane; the wavelet sample is read on the source row only
  int zs[SPB];
  unsigned srcb = 0u;
  const int kst = z0 - 4;
#pragma unroll
  for (int j = 0; j < SPB; ++j) {
    pw0[j] = pw1[j] = pw2[j] = pw3[j] = 0.0f;
    uw0[j] = uw1[j] = 0.0f;
    bw0[j] = bw1[j] = bw2[j] = bw3[j] = 0.0f;
    bu0[j] = bu1[j] = bu2[j] = bu3[j] = 0.0f;
    bpo[j] = 0.0f;
    qw0[j] = qw1[j] = qw2[j] = qw3[j] = 0.0f;
    const int s = s0 + min(j, ns - 1);
    zs[j] = j < ns ? d.src_z[s] : -1000;
    if (c == d.src_x[s]) srcb |= 1u << j;
    ww0[j] = bld(rW, vc, so[j] + rb(kst - 2));
    ww1[j] = bld(rW, vc, so[j] + rb(kst - 1));
    ww2[j] = bld(rW, vc, so[j] + rb(kst));
    cW[j] = bld(rW, vc, so[j] + rb(kst + 1));
    cU[j] = bld(rPU, vc1, so[j] + rb(kst));
    cP[j] = bld(rPU, vc, so[j] + rb(kst));
  }
  // coefficients: A at rows k (1-K1, A1) and k-2 (1-K2, A2, 1-K3); B at rows k-3 and k-5
  float ak1 = bld(rC, vc, rb(kst)), aa1 = bld(rC, vc, planeb + rb(kst));
  float ak2 = bld(rC, vc, 2u * planeb + rb(kst - 2)), aa2 = bld(rC, vc, 3u * planeb + rb(kst - 2));
  float ak3 = bld(rC, vc, 4u * planeb + rb(kst - 2));
  float bk1 = bld(rC, vc, rb(kst - 3)), ba1 = bld(rC, vc, planeb + rb(kst - 3));
  float bk2 = bld(rC, vc, 2u * planeb + rb(kst - 5)), ba2 = bld(rC, vc, 3u * planeb + rb(kst - 5));
  float bk3 = bld(rC, vc, 4u * planeb + rb(kst - 5));
  // energy of A's emitted rows, waiting for B's rows three iterations later
  float epa0 = 0.f, epa1 = 0.f, epa2 = 0.f, eua0 = 0.f, eua1 = 0.f, eua2 = 0.f;
  // old energy of row qq, loaded one iteration ahead (read-modify-write without a wait)
  const bool want_eu = want_u0 || want_u1;
  float ce0 = energy ? bld(rE, vc, eb + rb(kst - 5)) : 0.0f;
  float ce1 = (energy && want_eu) ? bld(rE, vc, eb + planeb + rb(kst - 5)) : 0.0f;

  const int iters = end_row - z0 + 9;   // k = z0-4 .. end_row+4
#pragma unroll 1
  for (int kk = 0; kk < iters; ++kk) {
    const int k = kst + kk;
    const int r = k - 2, q = k - 3, qq = k - 5;
    const unsigned rk = rb(k), rr = rb(r), rq = rb(q), rqq = rb(qq);
    // next row's loads before this row's stores
    float nW[SPB], nU[SPB], nP[SPB];
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      nW[j] = bld(rW, vc, so[j] + rb(k + 2));
      nU[j] = bld(rPU, vc1, so[j] + rb(k + 1));
      nP[j] = bld(rPU, vc, so[j] + rb(k + 1));
    }
    const float nak1 = bld(rC, vc, rb(k + 1)), naa1 = bld(rC, vc, planeb + rb(k + 1));
    const float nak2 = bld(rC, vc, 2u * planeb + rb(r + 1)), naa2 = bld(rC, vc, 3u * planeb + rb(r + 1));
    const float nak3 = bld(rC, vc, 4u * planeb + rb(r + 1));
    const float nbk1 = bld(rC, vc, rb(q + 1)), nba1 = bld(rC, vc, planeb + rb(q + 1));
    const float nbk2 = bld(rC, vc, 2u * planeb + rb(qq + 1)), nba2 = bld(rC, vc, 3u * planeb + rb(qq + 1));
    const float nbk3 = bld(rC, vc, 4u * planeb + rb(qq + 1));
    const float ne0 = energy ? bld(rE, vc, eb + rb(qq + 1)) : 0.0f;
    const float ne1 = (energy && want_eu) ? bld(rE, vc, eb + planeb + rb(qq + 1)) : 0.0f;

    const bool rowp = cp && k >= fs + 1 && k <= d.Z - 3;
    const bool rowu = cu && r >= fs && r <= d.Z - 2;
    const bool roww = cw && r >= fs && r <= d.Z - 3;
    const bool emit = kk >= 3;                                       // A's row r is complete
    const bool own_k = out_col && k >= z0 && k < end_row;
    const bool own_r = out_col && r >= z0 && r < end_row;
    const bool rowp2 = cp && q >= fs + 1 && q <= d.Z - 3;
    const bool rowu2 = cu && qq >= fs && qq <= d.Z - 2;
    const bool roww2 = cw && qq >= fs && qq <= d.Z - 3;
    const bool bstep = kk >= 6;                                      // B's p'' rows q-3..q are complete
    // Masked updates without branches: on the points a stage leaves alone its coefficients
    // become (1, 0), and 1*x - 0*d returns x (u is zeroed past column X-2, as the stage
    // leaves it).  On updated points the expression and its rounding are unchanged.
    const float pa = rowp ? ak1 : 1.0f, pb = rowp ? aa1 : 0.0f;
    const float ua = rowu ? ak2 : (u_ok ? 1.0f : 0.0f), ub = rowu ? aa2 : 0.0f;
    const float wa = roww ? ak3 : 1.0f, wb = roww ? aa2 : 0.0f;
    const float pa2 = rowp2 ? bk1 : 1.0f, pb2 = rowp2 ? ba1 : 0.0f;
    const float ua2 = rowu2 ? bk2 : (u_ok ? 1.0f : 0.0f), ub2 = rowu2 ? ba2 : 0.0f;
    const float wa2 = roww2 ? bk3 : 1.0f, wb2 = roww2 ? ba2 : 0.0f;
    const float dm = rowp ? 1.0f : 0.0f, dm2 = rowp2 ? 1.0f : 0.0f;   // stencil-image masks
    const bool own_q = out_col && q >= z0 && q < end_row;
    const bool own_qq = out_col && qq >= z0 && qq < end_row;
    // receiver slots of A's emitted row r (step t) and of B's final row qq (step t+1): the
    // lane's map entry is read only on rows with receivers (reading every row's entries one row
    // ahead with the other loads measured 7 % slower; the variants with records keep their
    // stores in branches, so a prefetched entry would only move the wait to the loop's end).
    // Owned rows come after the warm-up iterations: own_r implies emit, own_qq implies bstep.
    int ridx = -1, rfin = -1;
    if (recs && emit && rrows.has(r) && own_r) ridx = io.rmap[d.Z + r * d.pitch + c];
    if (recs && bstep && rrows.has(qq) && own_qq) rfin = io.rmap[d.Z + qq * d.pitch + c];
    float epa = 0.0f, eua = 0.0f, epb = 0.0f, eub = 0.0f;
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      // BF: every shot slot runs (a last group's slots past ns repeat its last shot's rows and
      // store nothing), so that no path through the loop skips the stores
      const bool jv = j == 0 || j < ns;
      if (BF || j < ns) {
        // ---- step A: new pressure at row k
        const float w3 = cW[j];
        float pn = cP[j];
        const float u0 = cU[j];
        const float um2 = shfl(u0, lane - 2), um1 = shfl(u0, lane - 1), up1 = shfl(u0, lane + 1);
        const float dd = C1 * (((u0 - um1) + ww2[j]) - ww1[j]) + C2 * (((up1 - um2) + w3) - ww0[j]);
        pn = pa * pn - pb * dd;
        if (k == zs[j]) {
          const float sv = d.src[(long long)(s0 + j) * d.nt + t];
          if ((srcb >> j) & 1u) pn = pn + sv;
        }
        pw0[j] = pw1[j]; pw1[j] = pw2[j]; pw2[j] = pw3[j]; pw3[j] = pn;
        if (d.fsurf && k == fs + 1) pw1[j] = -pn;
        if constexpr (FD) put(own_k && jv, dd * dm, rI, vs, so[j] + rk);
        // ---- step A: u', w' at row r; emit (p', u', w') of row r
        const float pr = pw1[j];
        const float pl1 = shfl(pr, lane - 1), pr1 = shfl(pr, lane + 1), pr2 = shfl(pr, lane + 2);
        const float un = ua * uw0[j] - ub * (C1 * (pr1 - pr) + C2 * (pr2 - pl1));
        const float wn = wa * ww0[j] - wb * (C1 * (pw2[j] - pr) + C2 * (pw3[j] - pw0[j]));
        if constexpr (FP) put(own_r && jv, pr, rJ, vs, so[j] + rr);
        epa += jv ? pr * pr : 0.0f;         // stored for owned points only
        eua += jv ? un * un : 0.0f;
        if (ridx >= 0 && jv) {
          const long long ro = ((long long)(s0 + j) * d.nt + t) * d.nrec + ridx;
          rec_p[ro] = pr;
          io.rec_u[ro] = un;
          io.rec_w[ro] = (d.fsurf && r == fs - 1) ? 0.0f : wn;   // row fs-1 is not sampled here
        }
        ww0[j] = ww1[j]; ww1[j] = ww2[j]; ww2[j] = w3;
        uw0[j] = uw1[j]; uw1[j] = u0;
        // ---- step B: window of A's rows (r = k-2)
        bw0[j] = bw1[j]; bw1[j] = bw2[j]; bw2[j] = bw3[j]; bw3[j] = wn;
        if (d.fsurf && r == fs) bw2[j] = wn;                 // w'(fs-1) = w'(fs)
        bu0[j] = bu1[j]; bu1[j] = bu2[j]; bu2[j] = bu3[j]; bu3[j] = un;
        const float p_old = bpo[j];                          // p'(q), q = r-1
        bpo[j] = pr;
        // ---- step B: new pressure at row q = k-3
        {
          const float v0 = bu2[j];                           // u'(q)
          const float vm2 = shfl(v0, lane - 2), vm1 = shfl(v0, lane - 1), vp1 = shfl(v0, lane + 1);
          const float dd2 = C1 * (((v0 - vm1) + bw2[j]) - bw1[j]) + C2 * (((vp1 - vm2) + bw3[j]) - bw0[j]);
          float pn2 = pa2 * p_old - pb2 * dd2;
          if (q == zs[j]) {
            const float sv = d.src[(long long)(s0 + j) * d.nt + t + 1];
            if ((srcb >> j) & 1u) pn2 = pn2 + sv;
          }
          qw0[j] = qw1[j]; qw1[j] = qw2[j]; qw2[j] = qw3[j]; qw3[j] = pn2;
          if (d.fsurf && q == fs + 1) qw1[j] = -pn2;
          if constexpr (FD) put(own_q && jv, dd2 * dm2, rI, vs1, so[j] + rq);
        }
        // ---- step B: u'', w'' at row qq = k-5; store the final rows
        {
          const float ps = qw1[j];
          const float sl1 = shfl(ps, lane - 1), sr1 = shfl(ps, lane + 1), sr2 = shfl(ps, lane + 2);
          const float un2 = ua2 * bu0[j] - ub2 * (C1 * (sr1 - ps) + C2 * (sr2 - sl1));
          const float wn2 = wa2 * bw0[j] - wb2 * (C1 * (qw2[j] - ps) + C2 * (qw3[j] - qw0[j]));
          if constexpr (BF) {
            const bool oq = own_qq && jv;
            bst(ps, oPU, oq ? vs : DROP, so[j] + rqq);
            if constexpr (FP) bst(ps, rJ, oq ? vs1 : DROP, so[j] + rqq);
            bst(un2, oPU, (oq && u_ok) ? vs1 : DROP, so[j] + rqq);
            bst(wn2, oW, (oq && qq < d.Z - 1 && !(d.fsurf && qq == fs - 1)) ? vs : DROP, so[j] + rqq);
            if (d.fsurf && qq == fs && oq) bst(wn2, oW, vs, so[j] + rqq - rowb);
          } else if (bstep && own_qq) {
            bst(ps, oPU, vs, so[j] + rqq);
            if (img_p1) bst(ps, rJ, vs1, so[j] + rqq);
            if (u_ok) bst(un2, oPU, vs1, so[j] + rqq);
            if (qq < d.Z - 1 && !(d.fsurf && qq == fs - 1)) bst(wn2, oW, vs, so[j] + rqq);
            if (d.fsurf && qq == fs) bst(wn2, oW, vs, so[j] + rqq - rowb);
          }
          if (rfin >= 0 && jv) {   // records of step t+1: the values just stored (points never written read zero)
            const long long ro = ((long long)(s0 + j) * d.nt + t + 1) * d.nrec + rfin;
            rec_p[ro] = ps;
            io.rec_u[ro] = u_ok ? un2 : 0.0f;
            io.rec_w[ro] = (qq < d.Z - 1 && !(d.fsurf && qq == fs - 1)) ? wn2 : 0.0f;
          }
          epb += jv ? ps * ps : 0.0f;       // stored for owned points only
          eub += jv ? un2 * un2 : 0.0f;
        }
      }
      cW[j] = nW[j]; cU[j] = nU[j]; cP[j] = nP[j];
    }
    ak1 = nak1; aa1 = naa1; ak2 = nak2; aa2 = naa2; ak3 = nak3;
    bk1 = nbk1; ba1 = nba1; bk2 = nbk2; ba2 = nba2; bk3 = nbk3;
    // energy: A's row r waits three iterations for B's row qq = r-3
    if (energy && bstep && own_qq) {
      const float ep = epa0 + epb;
      bst(ce0 + ep, rE, vs, eb + rqq);
      const float eu = (want_u0 ? eua0 : 0.0f) + (want_u1 ? eub : 0.0f);
      if (want_eu) bst(ce1 + eu, rE, vs, eb + planeb + rqq);
    }
    ce0 = ne0; ce1 = ne1;
    epa0 = epa1; epa1 = epa2; epa2 = epa;
    eua0 = eua1; eua1 = eua2; eua2 = eua;
  }
}

// ---------------------------------------------------------------------------
// adjoint step: transpose of stages 6..1 of step t, plus zero-lag imaging.
// Input: adjoint of the state after step t (receiver sources of step t added).
// Output: adjoint of the state after step t-1 (sources of t-1 not yet added).
// ---------------------------------------------------------------------------
template <int SPB, int UNR>
__global__ __launch_bounds__(NT) void ac_adj_step(
    AcDev d, const float *__restrict__ lin, float *__restrict__ lout, int t,
    const float *__restrict__ img_d, const float *__restrict__ img_p,
    float *__restrict__ grad_a1, float *__restrict__ grad_a2, float *__restrict__ grad_src) {
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - HALO_L + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= HALO_L && lane < HALO_L + SX && c < d.X;
  const bool cp = c >= 2 && c <= d.X - 3;
  const bool cu = c >= 1 && c <= d.X - 3;
  const bool cw = c >= 1 && c <= d.X - 2;
  const unsigned rowb = 4u * (unsigned)d.pitch;
  const unsigned planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  const int s0 = blockIdx.y * SPB;
  const int ns = min(SPB, d.S - s0);
  const int zmax = d.Z - 1;
  const int fs = d.fs;
  const int fsr = d.fsurf ? fs : -100;   // row whose adjoint collects the free-surface copy
  const bool want_a1 = grad_a1 != nullptr && img_d != nullptr;
  const bool want_a2 = grad_a2 != nullptr && img_p != nullptr;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rP = rsrc(lin, fieldb), rU = rsrc(lin + d.field, fieldb), rW = rsrc(lin + 2 * d.field, fieldb);
  const rsrc_t oP = rsrc(lout, fieldb), oU = rsrc(lout + d.field, fieldb), oW = rsrc(lout + 2 * d.field, fieldb);
  const rsrc_t rC = rsrc(coef_of(d, s0), 5u * planeb);
  const rsrc_t rI = rsrc(img_d, img_d ? fieldb : 0u), rIp = rsrc(img_p, img_p ? fieldb : 0u);
  const rsrc_t rG1 = rsrc(grad_a1, grad_a1 ? planeb * gridDim.y : 0u);
  const rsrc_t rG2 = rsrc(grad_a2, grad_a2 ? planeb * gridDim.y : 0u);
  const unsigned gb = planeb * blockIdx.y;
  unsigned so[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) so[j] = (unsigned)(s0 + min(j, ns - 1)) * planeb;

  // lambda_w' of row kr from its raw value: stage 6 transposed (row fs gathers row fs-1, then zero)
  auto lwfix = [&](unsigned soj, int kr, float raw) -> float {
    if (kr == fsr) return raw + bld(rW, vc, soj + rb(kr - 1));
    if (kr == fsr - 1) return 0.0f;
    return raw;
  };
  auto wmask = [&](int kr) -> bool { return cw && kr >= fs && kr <= d.Z - 3; };

  // shot-independent A2 window on w rows k-2..k+1 (masked)
  float aw0 = 0.0f;
  float aw1 = wmask(z0 - 3) ? bld(rC, vc, 3u * planeb + rb(z0 - 3)) : 0.0f;
  float aw2 = wmask(z0 - 2) ? bld(rC, vc, 3u * planeb + rb(z0 - 2)) : 0.0f;
  float aw3 = wmask(z0 - 1) ? bld(rC, vc, 3u * planeb + rb(z0 - 1)) : 0.0f;
  // per-shot windows: lw* = lambda_w' rows k-2..k+1; l2* = lambda_p2 rows k-2..k;
  // g* = A1 mask lambda_p2 rows k-3..k; lu* = lambda_u rows k-2..k; pp* = pressure image rows k-3..k
  float lw0[SPB], lw1[SPB], lw2[SPB], lw3[SPB];
  float l2a[SPB], l2b[SPB], l2c[SPB];
  float g0[SPB], g1[SPB], g2[SPB], g3[SPB];
  float lu0[SPB], lu1[SPB], lu2[SPB];
  float pp0[SPB], pp1[SPB], pp2[SPB], pp3[SPB];
  // row values loaded one iteration ahead: raw lambda_w(k+1), lambda_u(k), lambda_p(k), images(k)
  float cW[SPB], cU[SPB], cP[SPB], cI[SPB], cIp[SPB];
  int zs[SPB], xs[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) {
    lw0[j] = 0.0f;
    l2a[j] = l2b[j] = l2c[j] = 0.0f;
    g0[j] = g1[j] = g2[j] = g3[j] = 0.0f;
    lu0[j] = lu1[j] = lu2[j] = 0.0f;
    pp0[j] = pp1[j] = pp2[j] = pp3[j] = 0.0f;
    const int s = s0 + min(j, ns - 1);
    zs[j] = j < ns ? d.src_z[s] : -1000;
    xs[j] = d.src_x[s];
    lw1[j] = lwfix(so[j], z0 - 3, bld(rW, vc, so[j] + rb(z0 - 3)));
    lw2[j] = lwfix(so[j], z0 - 2, bld(rW, vc, so[j] + rb(z0 - 2)));
    lw3[j] = lwfix(so[j], z0 - 1, bld(rW, vc, so[j] + rb(z0 - 1)));
    cW[j] = bld(rW, vc, so[j] + rb(z0));
    cU[j] = bld(rU, vc, so[j] + rb(z0 - 1));
    cP[j] = bld(rP, vc, so[j] + rb(z0 - 1));
    cI[j] = want_a1 ? bld(rI, vc, so[j] + rb(z0 - 1)) : 0.0f;
    cIp[j] = want_a2 ? bld(rIp, vc, so[j] + rb(z0 - 1)) : 0.0f;
  }
  float cA2w = wmask(z0) ? bld(rC, vc, 3u * planeb + rb(z0)) : 0.0f;      // A2 at row k+1 (w window)
  float cA2u = bld(rC, vc, 3u * planeb + rb(z0 - 1));                   // A2 at row k (u)
  float cA1 = bld(rC, vc, planeb + rb(z0 - 1));
  float cK1 = bld(rC, vc, rb(z0 - 3)), cK2 = bld(rC, vc, 2u * planeb + rb(z0 - 3));
  float cK3 = bld(rC, vc, 4u * planeb + rb(z0 - 3));
  // old gradient-plane values of rows k (A1) and r (A2), loaded one iteration ahead so that the
  // read-modify-write waits for no load issued in the same iteration
  float cG1 = want_a1 ? bld(rG1, vc, gb + rb(z0 - 1)) : 0.0f;
  float cG2 = want_a2 ? bld(rG2, vc, gb + rb(z0 - 3)) : 0.0f;

#pragma unroll UNR
  for (int kk = 0; kk < d.tz + 3; ++kk) {
    const int k = z0 - 1 + kk;
    const unsigned rk = rb(k);
    const int r = k - 2;
    const unsigned rr = rb(r);
    // issue the next row's loads (all shots) before this row's stores
    float nW[SPB], nU[SPB], nP[SPB], nI[SPB], nIp[SPB];
    const unsigned rn = rb(k + 1), rn1 = rb(k + 2), rrn = rb(r + 1);
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      nW[j] = bld(rW, vc, so[j] + rn1);
      nU[j] = bld(rU, vc, so[j] + rn);
      nP[j] = bld(rP, vc, so[j] + rn);
      nI[j] = want_a1 ? bld(rI, vc, so[j] + rn) : 0.0f;
      nIp[j] = want_a2 ? bld(rIp, vc, so[j] + rn) : 0.0f;
    }
    const float nA2w = wmask(k + 2) ? bld(rC, vc, 3u * planeb + rn1) : 0.0f;
    const float nA2u = bld(rC, vc, 3u * planeb + rn), nA1 = bld(rC, vc, planeb + rn);
    const float nK1 = bld(rC, vc, rrn), nK2 = bld(rC, vc, 2u * planeb + rrn), nK3 = bld(rC, vc, 4u * planeb + rrn);
    const float nG1 = want_a1 ? bld(rG1, vc, gb + rn) : 0.0f;
    const float nG2 = want_a2 ? bld(rG2, vc, gb + rrn) : 0.0f;

    const bool rowp = cp && k >= fs + 1 && k <= d.Z - 3;
    const bool rowu_k = cu && k >= fs && k <= d.Z - 2;
    const bool pmr = cp && r >= fs + 1 && r <= d.Z - 3;
    const bool um = cu && r >= fs && r <= d.Z - 2;
    const bool wm = cw && r >= fs && r <= d.Z - 3;
    const bool outr = kk >= 3;
    const bool own = out_col && kk >= 1 && kk <= d.tz && k <= zmax;
    const bool wout = out_col && outr && r <= zmax;
    aw0 = aw1; aw1 = aw2; aw2 = aw3; aw3 = cA2w;
    const float a2k = rowu_k ? cA2u : 0.0f;
    const float a1k = rowp ? cA1 : 0.0f;
    const float k1r = pmr ? cK1 : 1.0f;
    const float k2r = um ? cK2 : 1.0f;
    const float k3r = wm ? cK3 : 1.0f;
    float ga1 = 0.0f, ga2 = 0.0f;
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      if (j < ns) {
        lw0[j] = lw1[j]; lw1[j] = lw2[j]; lw2[j] = lw3[j]; lw3[j] = lwfix(so[j], k + 1, cW[j]);
        lu0[j] = lu1[j]; lu1[j] = lu2[j]; lu2[j] = cU[j];
        g0[j] = g1[j]; g1[j] = g2[j]; g2[j] = g3[j];
        l2a[j] = l2b[j]; l2b[j] = l2c[j];
        if (want_a2) { pp0[j] = pp1[j]; pp1[j] = pp2[j]; pp2[j] = pp3[j]; pp3[j] = cIp[j]; }
        // stages 5/4 transposed: lambda_p3(k)
        const float fu = a2k * lu2[j];
        const float fum1 = shfl(fu, lane - 1), fum2 = shfl(fu, lane - 2), fup1 = shfl(fu, lane + 1);
        const float dxt = C1 * (fum1 - fu) + C2 * (fum2 - fup1);
        const float dzt = C1 * (aw1 * lw1[j] - aw2 * lw2[j]) + C2 * (aw0 * lw0[j] - aw3 * lw3[j]);
        const float l3 = cP[j] - (dxt + dzt);
        // stage 3 transposed: lambda_p2[fs+1] = lambda_p3[fs+1] - lambda_p3[fs-1]; lambda_p2[fs-1] = 0
        float l2 = l3;
        if (d.fsurf && k == fs + 1) {
          l2 = l3 - l2a[j];
          l2a[j] = 0.0f;
        }
        l2c[j] = l2;
        // stage 2 transposed: d loss / d src[s, t]
        if (grad_src && own && k == zs[j] && c == xs[j]) grad_src[(long long)(s0 + j) * d.nt + t] = l2;
        g3[j] = a1k * l2;
        if (want_a1 && own && rowp) ga1 -= l2 * cI[j];
        if (outr) {
          // stage 1 transposed and the three outputs at row r = k-2
          const float gc = g1[j];
          const float gl1 = shfl(gc, lane - 1), gr1 = shfl(gc, lane + 1), gr2 = shfl(gc, lane + 2);
          const float op = k1r * l2a[j];
          const float ou = k2r * lu0[j] - (C1 * (gc - gr1) + C2 * (gl1 - gr2));
          const float ow = k3r * lw0[j] - (C1 * (gc - g2[j]) + C2 * (g0[j] - g3[j]));
          if (want_a2) {
            const float pc = pp1[j];
            const float pl = shfl(pc, lane - 1), pr1 = shfl(pc, lane + 1), pr2 = shfl(pc, lane + 2);
            if (um) ga2 -= lu0[j] * (C1 * (pr1 - pc) + C2 * (pr2 - pl));
            if (wm) ga2 -= lw0[j] * (C1 * (pp2[j] - pc) + C2 * (pp3[j] - pp0[j]));
          }
          if (wout) {
            bst(op, oP, vs, so[j] + rr);
            if (c < d.X - 1) bst(ou, oU, vs, so[j] + rr);
            if (r < d.Z - 1) bst(ow, oW, vs, so[j] + rr);
          }
        }
      }
      cW[j] = nW[j]; cU[j] = nU[j]; cP[j] = nP[j]; cI[j] = nI[j]; cIp[j] = nIp[j];
    }
    cA2w = nA2w; cA2u = nA2u; cA1 = nA1; cK1 = nK1; cK2 = nK2; cK3 = nK3;
    // image terms of row k (stencil image) and row r (pressure image) into the group planes
    if (want_a1 && own && rowp) bst(cG1 + ga1, rG1, vs, gb + rk);
    if (want_a2 && wout) bst(cG2 + ga2, rG2, vs, gb + rr);
    cG1 = nG1; cG2 = nG2;
  }
}

// ---------------------------------------------------------------------------
// two adjoint steps in one launch
// ---------------------------------------------------------------------------
// Step B reverses step t+1 exactly as ac_adj_step does (its input is the adjoint of the state
// after step t+1 with the receiver sources of t+1 injected) and emits the finished row r = k-2
// of the adjoint state after step t.  The receiver sources of step t are added to the emitted
// rows on every lane (through the receiver-index map), and step A reverses step t from those
// rows three rows behind: its lambda_p3 at row q = k-3 and its outputs at row k-5, the only
// rows stored.  B's image terms wait three rows in registers so that one read-modify-write
// adds both steps to the gradient planes.
struct AdjPairIO {
  const float *in;
  float *out;
  int t;                                 // step A; step B is t+1
  const float *img_d0, *img_d1;          // stencil images of steps t and t+1 (d1 = d0 + one field)
  const float *img_p0, *img_p1;          // pressure images (A2 gradient), p1 = p0 + one field
  float *grad_a1, *grad_a2, *grad_src;
  const int *rmap;
  const float *g_p, *g_u, *g_w;          // record adjoints [S][nt][R]; step t's are added here, and
                                         // step t+1's (adj_pair_adds_first)
};

template <int SPB, bool A2, bool FULL, bool DR>
__global__ __launch_bounds__(NT) void ac_adj_pair(AcDev d, AdjPairIO io) {
  int z0, x0;
  if (!strip_of2(d, z0, x0)) return;
  // the lean variant (FULL = false) has no wavelet gradient and pressure record adjoints only:
  // the pointers it does not use are null constants, so they hold no scalar registers
  float *const grad_src = FULL ? io.grad_src : nullptr;
  const float *const g_u = FULL ? io.g_u : nullptr, *const g_w = FULL ? io.g_w : nullptr;
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - HALO2 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= HALO2 && lane < HALO2 + SX2 && c < d.X;
  const bool cp = c >= 2 && c <= d.X - 3;
  const bool cu = c >= 1 && c <= d.X - 3;
  const bool cw = c >= 1 && c <= d.X - 2;
  const bool in_col = c >= 0 && c < d.X;
  const unsigned rowb = 4u * (unsigned)d.pitch;
  const unsigned planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  const int s0 = blockIdx.y * SPB;
  const int ns = min(SPB, d.S - s0);
  const int zmax = d.Z - 1;
  const int fs = d.fs;
  const int fsr = d.fsurf ? fs : -100;   // row whose adjoint collects the free-surface copy
  const int end_row = min(z0 + d.tz, d.Z);
  const int t = io.t;
  const bool want_a1 = io.grad_a1 != nullptr && io.img_d0 != nullptr;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  // one descriptor spans two consecutive fields (lambda_p and lambda_u; the images of steps t
  // and t+1), the second reached through the lane offset (vc1 / vs1)
  const unsigned vc1 = vc + fieldb, vs1 = vs + fieldb;
  const rsrc_t rPU = rsrc(io.in, 2u * fieldb), rW = rsrc(io.in + 2 * d.field, fieldb);
  const rsrc_t oPU = rsrc(io.out, 2u * fieldb), oW = rsrc(io.out + 2 * d.field, fieldb);
  const rsrc_t rC = rsrc(coef_of(d, s0), 5u * planeb);
  const rsrc_t rI = rsrc(io.img_d0, want_a1 ? 2u * fieldb : 0u), rJ = rsrc(io.img_p0, A2 ? 2u * fieldb : 0u);
  const rsrc_t rG1 = rsrc(io.grad_a1, want_a1 ? planeb * gridDim.y : 0u);
  const rsrc_t rG2 = rsrc(io.grad_a2, A2 ? planeb * gridDim.y : 0u);
  const unsigned gb = planeb * blockIdx.y;
  const bool recs = io.rmap != nullptr && d.nrec > 0;
  // rows z0-6 .. z0+121 that hold receivers: the strip reads rows z0-6 .. end_row+5
  const adfwi::RowMask rrows = adfwi::receiver_rows(recs ? io.rmap : nullptr, d.Z, z0 - 6, lane);
  // receiver index at this lane's point of a row (-1: none).  In the row loop the raw map
  // entries are read one iteration ahead with the other loads (slot_ld, then slot); on rows
  // without receivers the access is dropped (no traffic, reads 0)
  const rsrc_t rM = rsrc(recs ? io.rmap + d.Z : nullptr, recs ? planeb : 0u);
  auto slot_hit = [&](int row) { return recs && rrows.has(row) && row >= 0 && row < d.Z && in_col; };
  auto slot_ld = [&](int row) -> float { return bld(rM, slot_hit(row) ? 4u * (unsigned)c : DROP, rb(row)); };
  auto slot = [&](int row, float raw) -> int { return slot_hit(row) ? __builtin_bit_cast(int, raw) : -1; };
  auto slot_at = [&](int row) -> int { return slot_hit(row) ? io.rmap[d.Z + row * d.pitch + c] : -1; };
  // BF: stores issued on every lane (dropped where a lane owns no point) and slots read one
  // iteration ahead, in the variants without the A2 gradient (with it the registers would cost
  // a wave per SIMD; those keep their stores in branches and read a slot in its iteration) and
  // when the host allows dropped accesses (DR)
  constexpr bool BF = DR && !A2;
  // the record adjoints of step t+1 at receiver slot m, added to the incoming adjoint state as
  // its rows are taken over from the loads (the state itself is left as it was); the variant
  // with both the A2 gradient and every record adjoint leaves them to a separate launch, as the
  // additions would cost it a wave per SIMD (adj_pair_adds_first)
  auto add_tb = [&](const float *g, int j, int m, float v) -> float {
    if constexpr (A2 && FULL) return v;
    return (g && m >= 0) ? v + g[((long long)(s0 + min(j, ns - 1)) * d.nt + t + 1) * d.nrec + m] : v;
  };
  unsigned so[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) so[j] = (unsigned)(s0 + min(j, ns - 1)) * planeb;
  auto wmask = [&](int kr) -> bool { return cw && kr >= fs && kr <= d.Z - 3; };
  auto lwfix = [&](unsigned soj, int kr, float raw) -> float {   // step B: raw rows from memory
    if (kr == fsr) return raw + bld(rW, vc, soj + rb(kr - 1));
    if (kr == fsr - 1) return 0.0f;
    return raw;
  };
  const int kst = z0 - 4;
  static_assert(TZ_MAX + 12 <= 128, "a strip's rows fit one receiver row mask");

  // step B windows (rows of k): lw* lambda_w' rows k-2..k+1, l2* lambda_p2 rows k-2..k,
  // g* A1-masked lambda_p2 rows k-3..k, lu* lambda_u rows k-2..k, pp* pressure images k-3..k;
  // c* the values of row k (k+1 for w) loaded one iteration ahead
  float lw0[SPB], lw1[SPB], lw2[SPB], lw3[SPB], l2a[SPB], l2b[SPB], l2c[SPB];
  float g0[SPB], g1[SPB], g2[SPB], g3[SPB], lu0[SPB], lu1[SPB], lu2[SPB];
  float pp0[SPB], pp1[SPB], pp2[SPB], pp3[SPB];
  float cW[SPB], cU[SPB], cP[SPB], cI[SPB], cIp[SPB];
  // step A windows, the same over the rows of q = k-3, fed by B's emitted rows; eP, eU, eW hold
  // B's previous emitted row; dI, dIp the step-t images of row q
  float mw0[SPB], mw1[SPB], mw2[SPB], mw3[SPB], m2a[SPB], m2b[SPB], m2c[SPB];
  float h0[SPB], h1[SPB], h2[SPB], h3[SPB], mu0[SPB], mu1[SPB], mu2[SPB];
  float qp0[SPB], qp1[SPB], qp2[SPB], qp3[SPB];
  float eP[SPB], eU[SPB], eW[SPB], dI[SPB], dIp[SPB];
  int zs[SPB], xs[SPB];
#pragma unroll
  for (int j = 0; j < SPB; ++j) {
    lw0[j] = 0.0f; l2a[j] = l2b[j] = l2c[j] = 0.0f; g0[j] = g1[j] = g2[j] = g3[j] = 0.0f;
    lu0[j] = lu1[j] = lu2[j] = 0.0f; pp0[j] = pp1[j] = pp2[j] = pp3[j] = 0.0f;
    mw0[j] = mw1[j] = mw2[j] = mw3[j] = 0.0f; m2a[j] = m2b[j] = m2c[j] = 0.0f;
    h0[j] = h1[j] = h2[j] = h3[j] = 0.0f; mu0[j] = mu1[j] = mu2[j] = 0.0f;
    qp0[j] = qp1[j] = qp2[j] = qp3[j] = 0.0f; eP[j] = eU[j] = eW[j] = 0.0f;
    const int s = s0 + min(j, ns - 1);
    zs[j] = j < ns ? d.src_z[s] : -1000;
    xs[j] = d.src_x[s];
    lw1[j] = lwfix(so[j], kst - 2, add_tb(g_w, j, slot_at(kst - 2), bld(rW, vc, so[j] + rb(kst - 2))));
    lw2[j] = lwfix(so[j], kst - 1, add_tb(g_w, j, slot_at(kst - 1), bld(rW, vc, so[j] + rb(kst - 1))));
    lw3[j] = lwfix(so[j], kst, add_tb(g_w, j, slot_at(kst), bld(rW, vc, so[j] + rb(kst))));
    cW[j] = add_tb(g_w, j, slot_at(kst + 1), bld(rW, vc, so[j] + rb(kst + 1)));
    cU[j] = add_tb(g_u, j, slot_at(kst), bld(rPU, vc1, so[j] + rb(kst)));
    cP[j] = add_tb(io.g_p, j, slot_at(kst), bld(rPU, vc, so[j] + rb(kst)));
    cI[j] = want_a1 ? bld(rI, vc1, so[j] + rb(kst)) : 0.0f;
    cIp[j] = A2 ? bld(rJ, vc1, so[j] + rb(kst)) : 0.0f;
    dI[j] = want_a1 ? bld(rI, vc, so[j] + rb(kst - 3)) : 0.0f;
    dIp[j] = A2 ? bld(rJ, vc, so[j] + rb(kst - 3)) : 0.0f;
  }
  // shot-independent coefficient windows: A2 on the w rows (B: k-2..k+1, A: q-2..q+1), and per
  // row A2 (u), A1 at k / q and 1-K1, 1-K2, 1-K3 at k-2 / q-2, loaded one iteration ahead
  float aw0 = 0.0f;
  float aw1 = wmask(kst - 2) ? bld(rC, vc, 3u * planeb + rb(kst - 2)) : 0.0f;
  float aw2 = wmask(kst - 1) ? bld(rC, vc, 3u * planeb + rb(kst - 1)) : 0.0f;
  float aw3 = wmask(kst) ? bld(rC, vc, 3u * planeb + rb(kst)) : 0.0f;
  float cA2w = wmask(kst + 1) ? bld(rC, vc, 3u * planeb + rb(kst + 1)) : 0.0f;
  float cA2u = bld(rC, vc, 3u * planeb + rb(kst)), cA1 = bld(rC, vc, planeb + rb(kst));
  float cK1 = bld(rC, vc, rb(kst - 2)), cK2 = bld(rC, vc, 2u * planeb + rb(kst - 2));
  float cK3 = bld(rC, vc, 4u * planeb + rb(kst - 2));
  const int qst = kst - 3;
  float bw0 = 0.0f;
  float bw1 = wmask(qst - 2) ? bld(rC, vc, 3u * planeb + rb(qst - 2)) : 0.0f;
  float bw2 = wmask(qst - 1) ? bld(rC, vc, 3u * planeb + rb(qst - 1)) : 0.0f;
  float bw3 = wmask(qst) ? bld(rC, vc, 3u * planeb + rb(qst)) : 0.0f;
  float dA2w = wmask(qst + 1) ? bld(rC, vc, 3u * planeb + rb(qst + 1)) : 0.0f;
  float dA2u = bld(rC, vc, 3u * planeb + rb(qst)), dA1 = bld(rC, vc, planeb + rb(qst));
  float dK1 = bld(rC, vc, rb(qst - 2)), dK2 = bld(rC, vc, 2u * planeb + rb(qst - 2));
  float dK3 = bld(rC, vc, 4u * planeb + rb(qst - 2));
  // step A at iteration k uses the coefficient rows step B used at iteration k-3: without the A2
  // gradient B's values of the last two iterations wait in registers (h1: k-1, h2: k-2) instead
  // of being loaded again (with it the registers would cost a wave per SIMD, and A loads them)
  constexpr bool DELAY = !A2;
  auto bA2w = [&](int j) { return wmask(j + 1) ? bld(rC, vc, 3u * planeb + rb(j + 1)) : 0.0f; };
  float h1A2w = 0.f, h2A2w = 0.f, h1A2u = 0.f, h2A2u = 0.f, h1A1 = 0.f, h2A1 = 0.f;
  float h1K1 = 0.f, h2K1 = 0.f, h1K2 = 0.f, h2K2 = 0.f, h1K3 = 0.f, h2K3 = 0.f;
  if constexpr (DELAY) {
    h1A2w = bA2w(kst - 1); h2A2w = bA2w(kst - 2);
    h1A2u = bld(rC, vc, 3u * planeb + rb(kst - 1)); h2A2u = bld(rC, vc, 3u * planeb + rb(kst - 2));
    h1A1 = bld(rC, vc, planeb + rb(kst - 1)); h2A1 = bld(rC, vc, planeb + rb(kst - 2));
    h1K1 = bld(rC, vc, rb(kst - 3)); h2K1 = bld(rC, vc, rb(kst - 4));
    h1K2 = bld(rC, vc, 2u * planeb + rb(kst - 3)); h2K2 = bld(rC, vc, 2u * planeb + rb(kst - 4));
    h1K3 = bld(rC, vc, 4u * planeb + rb(kst - 3)); h2K3 = bld(rC, vc, 4u * planeb + rb(kst - 4));
  }
  // B's image terms of rows k and r waiting three iterations for A's rows q and qq
  float ga1d0 = 0.f, ga1d1 = 0.f, ga1d2 = 0.f, ga2d0 = 0.f, ga2d1 = 0.f, ga2d2 = 0.f;
  // old gradient-plane values of rows q (A1) and qq (A2), loaded one iteration ahead
  float cG1 = want_a1 ? bld(rG1, vc, gb + rb(qst)) : 0.0f;
  float cG2 = A2 ? bld(rG2, vc, gb + rb(kst - 5)) : 0.0f;
  // raw map entries of B's emitted row r and of the incoming rows k+1 (p, u) and k+2 (w)
  float cR = 0.f, cMP = 0.f, cMW = 0.f;
  if constexpr (BF) { cR = slot_ld(kst - 2); cMP = slot_ld(kst + 1); cMW = slot_ld(kst + 2); }

  const int iters = end_row - z0 + 9;   // k = z0-4 .. end_row+4
#pragma unroll 1
  for (int kk = 0; kk < iters; ++kk) {
    const int k = kst + kk;
    const int r = k - 2, q = k - 3, qq = k - 5;
    const unsigned rk = rb(k), rr = rb(r), rq = rb(q), rqq = rb(qq);
    // next row's loads before this row's stores
    float nR = 0.f, nMW = 0.f;
    if constexpr (BF) { nR = slot_ld(r + 1); nMW = slot_ld(k + 3); }
    float nW[SPB], nU[SPB], nP[SPB], nI[SPB], nIp[SPB], mI[SPB], mIp[SPB];
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      nW[j] = bld(rW, vc, so[j] + rb(k + 2));
      nU[j] = bld(rPU, vc1, so[j] + rb(k + 1));
      nP[j] = bld(rPU, vc, so[j] + rb(k + 1));
      nI[j] = want_a1 ? bld(rI, vc1, so[j] + rb(k + 1)) : 0.0f;
      nIp[j] = A2 ? bld(rJ, vc1, so[j] + rb(k + 1)) : 0.0f;
      mI[j] = want_a1 ? bld(rI, vc, so[j] + rb(q + 1)) : 0.0f;
      mIp[j] = A2 ? bld(rJ, vc, so[j] + rb(q + 1)) : 0.0f;
    }
    const float nA2w = wmask(k + 2) ? bld(rC, vc, 3u * planeb + rb(k + 2)) : 0.0f;
    const float nA2u = bld(rC, vc, 3u * planeb + rb(k + 1)), nA1 = bld(rC, vc, planeb + rb(k + 1));
    const float nK1 = bld(rC, vc, rb(r + 1)), nK2 = bld(rC, vc, 2u * planeb + rb(r + 1));
    const float nK3 = bld(rC, vc, 4u * planeb + rb(r + 1));
    float mA2w = 0.f, mA2u = 0.f, mA1 = 0.f, mK1 = 0.f, mK2 = 0.f, mK3 = 0.f;
    if constexpr (!DELAY) {
      mA2w = wmask(q + 2) ? bld(rC, vc, 3u * planeb + rb(q + 2)) : 0.0f;
      mA2u = bld(rC, vc, 3u * planeb + rb(q + 1)); mA1 = bld(rC, vc, planeb + rb(q + 1));
      mK1 = bld(rC, vc, rb(qq + 1)); mK2 = bld(rC, vc, 2u * planeb + rb(qq + 1));
      mK3 = bld(rC, vc, 4u * planeb + rb(qq + 1));
    }
    const float nG1 = want_a1 ? bld(rG1, vc, gb + rb(q + 1)) : 0.0f;
    const float nG2 = A2 ? bld(rG2, vc, gb + rb(qq + 1)) : 0.0f;

    // step B masks (rows k, r) and step A masks (rows q, qq)
    const bool rowp = cp && k >= fs + 1 && k <= d.Z - 3;
    const bool rowu_k = cu && k >= fs && k <= d.Z - 2;
    const bool pmr = cp && r >= fs + 1 && r <= d.Z - 3;
    const bool um = cu && r >= fs && r <= d.Z - 2;
    const bool wm = cw && r >= fs && r <= d.Z - 3;
    const bool rowp_q = cp && q >= fs + 1 && q <= d.Z - 3;
    const bool rowu_q = cu && q >= fs && q <= d.Z - 2;
    const bool pmr_q = cp && qq >= fs + 1 && qq <= d.Z - 3;
    const bool um_q = cu && qq >= fs && qq <= d.Z - 2;
    const bool wm_q = cw && qq >= fs && qq <= d.Z - 3;
    const bool own_k = out_col && k >= z0 && k < end_row;
    const bool own_r = out_col && r >= z0 && r < end_row;
    const bool own_q = out_col && q >= z0 && q < end_row;
    const bool own_qq = out_col && qq >= z0 && qq < end_row;
    aw0 = aw1; aw1 = aw2; aw2 = aw3; aw3 = cA2w;
    bw0 = bw1; bw1 = bw2; bw2 = bw3; bw3 = dA2w;
    const float a2k = rowu_k ? cA2u : 0.0f, a1k = rowp ? cA1 : 0.0f;
    const float k1r = pmr ? cK1 : 1.0f, k2r = um ? cK2 : 1.0f, k3r = wm ? cK3 : 1.0f;
    const float a2q = rowu_q ? dA2u : 0.0f, a1q = rowp_q ? dA1 : 0.0f;
    const float k1q = pmr_q ? dK1 : 1.0f, k2q = um_q ? dK2 : 1.0f, k3q = wm_q ? dK3 : 1.0f;
    // A1 image-term masks as 0/1 factors on one operand: (x*1)*y rounds as x*y, and a masked
    // term is +-0, which leaves the sum unchanged; the sums are stored for owned points only
    const float m1k = rowp ? 1.0f : 0.0f, m1q = rowp_q ? 1.0f : 0.0f;
    // receiver slots: B's emitted row r (step t's sources, on every lane of the strip), and the
    // rows of the incoming state B reads now, p and u of row k and w of row k+1 (step t+1's)
    int ridx, mpn, mwn;
    if constexpr (BF) { ridx = slot(r, cR); mpn = slot(k + 1, cMP); mwn = slot(k + 2, cMW); }
    else { ridx = slot_at(r); mpn = slot_at(k + 1); mwn = slot_at(k + 2); }
    float ga1b = 0.0f, ga2b = 0.0f, ga1a = 0.0f, ga2a = 0.0f;
#pragma unroll
    for (int j = 0; j < SPB; ++j) {
      // BF: every shot slot runs (a last group's slots past ns repeat its last shot's rows and
      // store nothing), so that no path through the loop skips the stores
      const bool jv = j == 0 || j < ns;
      if (BF || j < ns) {
        // ================= step B (t+1), as ac_adj_step at row k
        lw0[j] = lw1[j]; lw1[j] = lw2[j]; lw2[j] = lw3[j]; lw3[j] = lwfix(so[j], k + 1, cW[j]);
        lu0[j] = lu1[j]; lu1[j] = lu2[j]; lu2[j] = cU[j];
        g0[j] = g1[j]; g1[j] = g2[j]; g2[j] = g3[j];
        l2a[j] = l2b[j]; l2b[j] = l2c[j];
        if (A2) { pp0[j] = pp1[j]; pp1[j] = pp2[j]; pp2[j] = pp3[j]; pp3[j] = cIp[j]; }
        const float fu = a2k * lu2[j];
        const float fum1 = shfl(fu, lane - 1), fum2 = shfl(fu, lane - 2), fup1 = shfl(fu, lane + 1);
        const float dxt = C1 * (fum1 - fu) + C2 * (fum2 - fup1);
        const float dzt = C1 * (aw1 * lw1[j] - aw2 * lw2[j]) + C2 * (aw0 * lw0[j] - aw3 * lw3[j]);
        const float l3 = cP[j] - (dxt + dzt);
        float l2 = l3;
        if (d.fsurf && k == fs + 1) { l2 = l3 - l2a[j]; l2a[j] = 0.0f; }
        l2c[j] = l2;
        if (grad_src && own_k && k == zs[j] && c == xs[j]) grad_src[(long long)(s0 + j) * d.nt + t + 1] = l2;
        g3[j] = a1k * l2;
        if (want_a1) ga1b -= jv ? l2 * (cI[j] * m1k) : 0.0f;
        // outputs of B at row r: the adjoint state after step t (plus step t's sources)
        const float gc = g1[j];
        const float gl1 = shfl(gc, lane - 1), gr1 = shfl(gc, lane + 1), gr2 = shfl(gc, lane + 2);
        float op = k1r * l2a[j];
        float ou = k2r * lu0[j] - (C1 * (gc - gr1) + C2 * (gl1 - gr2));
        float ow = k3r * lw0[j] - (C1 * (gc - g2[j]) + C2 * (g0[j] - g3[j]));
        if (A2) {
          const float pc = pp1[j];
          const float pl = shfl(pc, lane - 1), pr1 = shfl(pc, lane + 1), pr2 = shfl(pc, lane + 2);
          if (own_r) {
            if (um) ga2b -= lu0[j] * (C1 * (pr1 - pc) + C2 * (pr2 - pl));
            if (wm) ga2b -= lw0[j] * (C1 * (pp2[j] - pc) + C2 * (pp3[j] - pp0[j]));
          }
        }
        if (c >= d.X - 1) ou = 0.0f;          // rows/columns the one-step kernel never writes
        if (r >= d.Z - 1) ow = 0.0f;
        if (ridx >= 0 && jv) {
          const long long gi = ((long long)(s0 + j) * d.nt + t) * d.nrec + ridx;
          if (io.g_p) op = op + io.g_p[gi];
          if (g_u) ou = ou + g_u[gi];
          if (g_w) ow = ow + g_w[gi];
        }
        // ================= step A (t) at row q = k-3 over B's rows
        const float wprev = eW[j];
        mw0[j] = mw1[j]; mw1[j] = mw2[j]; mw2[j] = mw3[j];
        mw3[j] = (q + 1 == fsr) ? ow + wprev : ((q + 1 == fsr - 1) ? 0.0f : ow);
        mu0[j] = mu1[j]; mu1[j] = mu2[j]; mu2[j] = eU[j];
        h0[j] = h1[j]; h1[j] = h2[j]; h2[j] = h3[j];
        m2a[j] = m2b[j]; m2b[j] = m2c[j];
        if (A2) { qp0[j] = qp1[j]; qp1[j] = qp2[j]; qp2[j] = qp3[j]; qp3[j] = dIp[j]; }
        const float fv = a2q * mu2[j];
        const float fvm1 = shfl(fv, lane - 1), fvm2 = shfl(fv, lane - 2), fvp1 = shfl(fv, lane + 1);
        const float ext = C1 * (fvm1 - fv) + C2 * (fvm2 - fvp1);
        const float ezt = C1 * (bw1 * mw1[j] - bw2 * mw2[j]) + C2 * (bw0 * mw0[j] - bw3 * mw3[j]);
        const float m3 = eP[j] - (ext + ezt);
        float m2 = m3;
        if (d.fsurf && q == fs + 1) { m2 = m3 - m2a[j]; m2a[j] = 0.0f; }
        m2c[j] = m2;
        if (grad_src && own_q && q == zs[j] && c == xs[j]) grad_src[(long long)(s0 + j) * d.nt + t] = m2;
        h3[j] = a1q * m2;
        if (want_a1) ga1a -= jv ? m2 * (dI[j] * m1q) : 0.0f;
        {
          const float hc = h1[j];
          const float hl1 = shfl(hc, lane - 1), hr1 = shfl(hc, lane + 1), hr2 = shfl(hc, lane + 2);
          const float opa = k1q * m2a[j];
          const float oua = k2q * mu0[j] - (C1 * (hc - hr1) + C2 * (hl1 - hr2));
          const float owa = k3q * mw0[j] - (C1 * (hc - h2[j]) + C2 * (h0[j] - h3[j]));
          if (A2) {
            const float pc = qp1[j];
            const float pl = shfl(pc, lane - 1), pr1 = shfl(pc, lane + 1), pr2 = shfl(pc, lane + 2);
            if (own_qq) {
              if (um_q) ga2a -= mu0[j] * (C1 * (pr1 - pc) + C2 * (pr2 - pl));
              if (wm_q) ga2a -= mw0[j] * (C1 * (qp2[j] - pc) + C2 * (qp3[j] - qp0[j]));
            }
          }
          if constexpr (BF) {
            const bool oq = own_qq && jv;
            bst(opa, oPU, oq ? vs : DROP, so[j] + rqq);
            bst(oua, oPU, (oq && c < d.X - 1) ? vs1 : DROP, so[j] + rqq);
            bst(owa, oW, (oq && qq < d.Z - 1) ? vs : DROP, so[j] + rqq);
          } else if (own_qq) {
            bst(opa, oPU, vs, so[j] + rqq);
            if (c < d.X - 1) bst(oua, oPU, vs1, so[j] + rqq);
            if (qq < d.Z - 1) bst(owa, oW, vs, so[j] + rqq);
          }
        }
        eP[j] = op; eU[j] = ou; eW[j] = ow;
      }
      cW[j] = add_tb(g_w, j, mwn, nW[j]); cU[j] = add_tb(g_u, j, mpn, nU[j]); cP[j] = add_tb(io.g_p, j, mpn, nP[j]);
      cI[j] = nI[j]; cIp[j] = nIp[j];
      dI[j] = mI[j]; dIp[j] = mIp[j];
    }
    if constexpr (DELAY) {
      dA2w = h2A2w; dA2u = h2A2u; dA1 = h2A1; dK1 = h2K1; dK2 = h2K2; dK3 = h2K3;   // rows of iteration k-2
      h2A2w = h1A2w; h2A2u = h1A2u; h2A1 = h1A1; h2K1 = h1K1; h2K2 = h1K2; h2K3 = h1K3;
      h1A2w = cA2w; h1A2u = cA2u; h1A1 = cA1; h1K1 = cK1; h1K2 = cK2; h1K3 = cK3;
    } else {
      dA2w = mA2w; dA2u = mA2u; dA1 = mA1; dK1 = mK1; dK2 = mK2; dK3 = mK3;
    }
    cA2w = nA2w; cA2u = nA2u; cA1 = nA1; cK1 = nK1; cK2 = nK2; cK3 = nK3;
    cR = nR; cMP = cMW; cMW = nMW;
    // image terms: B's row k waits three iterations for A's row q, B's row r for A's row qq
    // (without the A1 gradient rG1 is empty and a BF store is dropped)
    if constexpr (BF) bst(cG1 + (ga1d0 + ga1a), rG1, own_q ? vs : DROP, gb + rq);
    else if (want_a1 && own_q) bst(cG1 + (ga1d0 + ga1a), rG1, vs, gb + rq);
    if (A2 && own_qq) bst(cG2 + (ga2d0 + ga2a), rG2, vs, gb + rqq);
    cG1 = nG1; cG2 = nG2;
    ga1d0 = ga1d1; ga1d1 = ga1d2; ga1d2 = ga1b;
    ga2d0 = ga2d1; ga2d1 = ga2d2; ga2d2 = ga2b;
  }
}

// rec_{p,u,w}[s][t][r] = state at the receivers after step t
__global__ void ac_gather(AcDev d, const float *__restrict__ st, float *__restrict__ rp, float *__restrict__ ru,
                          float *__restrict__ rw, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  t += time_base(d);
  if (t < 0 || t >= d.nt) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const long long o = (long long)s * d.plane + (long long)d.rcv_z[r] * d.pitch + d.rcv_x[r];
  const long long ro = ((long long)s * d.nt + t) * d.nrec + r;
  rp[ro] = st[o];
  ru[ro] = st[d.field + o];
  rw[ro] = st[2 * d.field + o];
}

// lam[field][s][zr][xr] += g[s][t][r]: receiver adjoint sources of step t
__global__ void ac_inject(AcDev d, float *__restrict__ lam, const float *__restrict__ g_p,
                          const float *__restrict__ g_u, const float *__restrict__ g_w, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const long long o = (long long)s * d.plane + (long long)d.rcv_z[r] * d.pitch + d.rcv_x[r];
  const long long gi = ((long long)s * d.nt + t) * d.nrec + r;
  if (g_p) atomicAdd(&lam[o], g_p[gi]);
  if (g_u) atomicAdd(&lam[d.field + o], g_u[gi]);
  if (g_w) atomicAdd(&lam[2 * d.field + o], g_w[gi]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int layout_of(const adfwi_acoustic_desc *d, adfwi_layout *L) {
  if (!d) { adfwi::set_error("null descriptor"); return ADFWI_E_ARG; }
  if (d->abi_version != ADFWI_ABI_VERSION || d->struct_size != (int32_t)sizeof(adfwi_acoustic_desc)) {
    adfwi::set_error("acoustic descriptor version %d size %d, library expects %d / %d", d->abi_version,
                     d->struct_size, ADFWI_ABI_VERSION, (int)sizeof(adfwi_acoustic_desc));
    return ADFWI_E_VERSION;
  }
  ADFWI_CHECK_ARG(d->nshots >= 1, "nshots must be >= 1 (got %d)", d->nshots);
  ADFWI_CHECK_ARG(d->nz >= 8 && d->nx >= 8, "padded grid %d x %d too small (need >= 8 x 8)", d->nz, d->nx);
  ADFWI_CHECK_ARG(d->pitch >= d->nx, "pitch %d < nx %d", d->pitch, d->nx);
  ADFWI_CHECK_ARG(d->nt >= 1 && d->nrec >= 0, "nt %d / nrec %d invalid", d->nt, d->nrec);
  if (d->free_surface) {
    ADFWI_CHECK_ARG(d->fs >= 1 && d->fs + 3 <= d->nz - 3, "free-surface row %d outside grid of %d rows",
                    d->fs, d->nz);
  } else {
    ADFWI_CHECK_ARG(d->fs == 1, "fs must be 1 without a free surface (got %d)", d->fs);
  }
  L->zbase = d->free_surface ? d->fs - 1 : 0;
  L->ntiles_x = (d->nx + SX - 1) / SX;
  L->tile_x = SX;
  // shots per wavefront (1, 2, 4 or 8, a kernel template) and strip height (4..64 rows): the
  // largest shot group (<= 4) with 16-row strips that still gives >= 4096 wavefronts (about four
  // per SIMD on 256 CUs), then the tallest strips (32 down to 4 rows) that keep that count; the
  // two-step kernel uses strips twice as tall.  Measured on MI355X for a 64-shot 464 x 1232 grid:
  // 4 shots x 32 rows (64 for the two-step kernel) beats 2 shots and 8/16-row strips.
  const int rows = d->nz - L->zbase;
  auto waves = [&](int spb_, int tz_) {
    return (long long)L->ntiles_x * ((rows + tz_ - 1) / tz_) * ((d->nshots + spb_ - 1) / spb_);
  };
  int spb = d->shots_per_block;
  int tz = d->tile_z;
  ADFWI_CHECK_ARG(spb == 0 || spb == 1 || spb == 2 || spb == 4 || spb == 8,
                  "shots_per_block must be 0, 1, 2, 4 or 8 (got %d)", spb);
  ADFWI_CHECK_ARG(tz == 0 || (tz >= 4 && tz <= TZ_MAX), "tile_z must be 0 or 4..%d (got %d)", TZ_MAX, tz);
  if (spb == 0) {
    spb = 1;
    for (int cand = 4; cand >= 1; cand >>= 1)
      if (waves(cand, tz ? tz : 16) >= 4096) { spb = cand; break; }
  }
  if (tz == 0) {
    tz = 4;
    for (int cand = 32; cand >= 4; cand >>= 1)
      if (waves(spb, cand) >= 4096) { tz = cand; break; }
  }
  L->tile_z = tz;
  L->ntiles_z = (rows + tz - 1) / tz;
  L->ntiles = L->ntiles_z * L->ntiles_x;
  ADFWI_CHECK_ARG(d->row_unroll == 0 || d->row_unroll == 1 || d->row_unroll == 2 || d->row_unroll == 4,
                  "row_unroll must be 0, 1, 2 or 4 (got %d)", d->row_unroll);
  L->shots_per_block = spb;
  L->ngroups = (d->nshots + spb - 1) / spb;
  L->nfields = 3;
  // the two-step adjoint holds twice the windows of the one-step kernel: 2 shots per wavefront
  // keep it at 3 waves per SIMD where 4 shots would leave 2
  int aspb = d->adj_shots_per_block;
  ADFWI_CHECK_ARG(aspb == 0 || aspb == 1 || aspb == 2 || aspb == 4,
                  "adj_shots_per_block must be 0, 1, 2 or 4 (got %d)", aspb);
  if (aspb == 0) aspb = spb < 2 ? spb : 2;
  L->reserved = aspb;
  L->field_floats = (int64_t)d->nshots * d->nz * d->pitch;
  // buffer descriptors address one field (all shots) with 32-bit byte offsets
  ADFWI_CHECK_ARG(L->field_floats * 4 < ((int64_t)1 << 31),
                  "one field of %d shots x %d x %d exceeds 2 GiB; advance the shots in smaller launches",
                  d->nshots, d->nz, d->pitch);
  L->state_floats = 3 * L->field_floats;
  const int finer = spb < aspb ? spb : aspb;
  L->accum_floats = (int64_t)((d->nshots + finer - 1) / finer) * d->nz * d->pitch;
  return ADFWI_OK;
}

int check_device_desc(const adfwi_acoustic_desc *d, bool need_records) {
  ADFWI_CHECK_ARG(d->coef && d->src && d->src_z && d->src_x, "null coefficient/source pointer");
  if (need_records && d->nrec > 0) ADFWI_CHECK_ARG(d->rcv_z && d->rcv_x, "null receiver table pointer");
  return ADFWI_OK;
}

AcDev make_dev(const adfwi_acoustic_desc *d, const adfwi_layout &L) {
  AcDev v;
  v.S = d->nshots; v.Z = d->nz; v.X = d->nx; v.pitch = d->pitch;
  v.fs = d->fs; v.fsurf = d->free_surface ? 1 : 0; v.nt = d->nt; v.nrec = d->nrec;
  v.zbase = L.zbase; v.nsx = L.ntiles_x; v.nsz = L.ntiles_z; v.nstrips = L.ntiles; v.tz = L.tile_z;
  v.spb = L.shots_per_block;
  v.unroll = d->row_unroll > 0 ? d->row_unroll : 2;
  v.plane = (long long)d->nz * d->pitch;
  v.field = (long long)d->nshots * v.plane;
  v.coef = d->coef; v.src = d->src; v.src_z = d->src_z; v.src_x = d->src_x;
  v.coef_set = d->coef_set;
  v.rcv_z = d->rcv_z; v.rcv_x = d->rcv_x;
  v.tbase = nullptr;
  v.eu_from = d->nt;
  return v;
}

// the branch-free two-step variants drop the accesses a lane must not make by pointing them
// DROP bytes past the buffer: every descriptor they use (two fields at most) must span under
// 2 GiB.  Larger fields (and ADFWI_PAIR_DROP=0) take the variants with stores in branches.
bool pair_drop_ok(const adfwi_layout &L) {
  const char *e = std::getenv("ADFWI_PAIR_DROP");
  if (e && e[0] == '0') return false;
  return 2 * L.field_floats * 4 < ((int64_t)1 << 31);
}

dim3 step_grid(const adfwi_layout &L) { return dim3((L.ntiles + WPB - 1) / WPB, L.ngroups); }

template <int SPB>
void launch_fwd_u(const AcDev &v, dim3 g, dim3 b, hipStream_t st, const float *in, float *out, int t, float *e,
                  int acc_u, float *imd, float *imp) {
  switch (v.unroll) {
    case 4: hipLaunchKernelGGL((ac_fwd_step<SPB, 4>), g, b, 0, st, v, in, out, t, e, acc_u, imd, imp); break;
    case 2: hipLaunchKernelGGL((ac_fwd_step<SPB, 2>), g, b, 0, st, v, in, out, t, e, acc_u, imd, imp); break;
    default: hipLaunchKernelGGL((ac_fwd_step<SPB, 1>), g, b, 0, st, v, in, out, t, e, acc_u, imd, imp);
  }
}

void launch_fwd(const AcDev &v, dim3 g, dim3 b, hipStream_t st, const float *in, float *out, int t, float *e,
                int acc_u, float *imd, float *imp) {
  switch (v.spb) {
    case 8: launch_fwd_u<8>(v, g, b, st, in, out, t, e, acc_u, imd, imp); break;
    case 4: launch_fwd_u<4>(v, g, b, st, in, out, t, e, acc_u, imd, imp); break;
    case 2: launch_fwd_u<2>(v, g, b, st, in, out, t, e, acc_u, imd, imp); break;
    default: launch_fwd_u<1>(v, g, b, st, in, out, t, e, acc_u, imd, imp);
  }
}

// the pair kernel tiles the grid with 52-column strips of twice the single-step height
AcDev pair_dev(const AcDev &v) {
  AcDev w = v;
  w.tz = min(TZ_MAX, 2 * v.tz);
  w.nsx = (v.X + SX2 - 1) / SX2;
  w.nsz = (v.Z - v.zbase + w.tz - 1) / w.tz;
  w.nstrips = w.nsx * w.nsz;
  return w;
}

template <int F>
void launch_pair_f(const AcDev &w, dim3 g, dim3 b, hipStream_t st, const PairIO &io) {
  switch (w.spb) {
    case 4: hipLaunchKernelGGL((ac_fwd_pair<4, F>), g, b, 0, st, w, io); break;
    case 2: hipLaunchKernelGGL((ac_fwd_pair<2, F>), g, b, 0, st, w, io); break;
    default: hipLaunchKernelGGL((ac_fwd_pair<1, F>), g, b, 0, st, w, io);
  }
}

// the variant without the features this launch does not use (fewer live scalar registers)
void launch_pair(const AcDev &w, dim3 g, dim3 b, hipStream_t st, const PairIO &io, bool drop_ok) {
  const bool rec = io.energy || io.rec_p;
  const bool imd = io.img_d0 || io.img_d1, imp = io.img_p0 || io.img_p1;
  if (!imd && !imp) launch_pair_f<PF_REC>(w, g, b, st, io);
  else if (!rec && !imp)
    drop_ok ? launch_pair_f<PF_IMGD>(w, g, b, st, io) : launch_pair_f<PF_IMGD | PF_BRANCH>(w, g, b, st, io);
  else if (!rec)
    drop_ok ? launch_pair_f<PF_IMGD | PF_IMGP>(w, g, b, st, io)
            : launch_pair_f<PF_IMGD | PF_IMGP | PF_BRANCH>(w, g, b, st, io);
  else if (!imp) launch_pair_f<PF_REC | PF_IMGD>(w, g, b, st, io);
  else launch_pair_f<PF_REC | PF_IMGD | PF_IMGP>(w, g, b, st, io);
}

template <bool A2, bool FULL, bool DR>
void launch_adj_pair_a(const AcDev &w, dim3 g, dim3 b, hipStream_t st, const AdjPairIO &io) {
  switch (w.spb) {
    case 4: hipLaunchKernelGGL((ac_adj_pair<4, A2, FULL, DR>), g, b, 0, st, w, io); break;
    case 2: hipLaunchKernelGGL((ac_adj_pair<2, A2, FULL, DR>), g, b, 0, st, w, io); break;
    default: hipLaunchKernelGGL((ac_adj_pair<1, A2, FULL, DR>), g, b, 0, st, w, io);
  }
}

// whether the adjoint pair variant chosen for io adds the record adjoints of its first reversed
// step itself (all but the A2 + FULL variant)
bool adj_pair_adds_first(const AdjPairIO &io) {
  const bool full = io.grad_src || io.g_u || io.g_w;
  const bool a2 = io.grad_a2 && io.img_p0 && io.img_p1;
  return !(a2 && full);
}

void launch_adj_pair(const AcDev &w, dim3 g, dim3 b, hipStream_t st, const AdjPairIO &io, bool drop_ok) {
  const bool full = io.grad_src || io.g_u || io.g_w;
  const bool a2 = io.grad_a2 && io.img_p0 && io.img_p1;
  if (a2)
    full ? launch_adj_pair_a<true, true, false>(w, g, b, st, io) : launch_adj_pair_a<true, false, false>(w, g, b, st, io);
  else if (drop_ok)
    full ? launch_adj_pair_a<false, true, true>(w, g, b, st, io) : launch_adj_pair_a<false, false, true>(w, g, b, st, io);
  else
    full ? launch_adj_pair_a<false, true, false>(w, g, b, st, io) : launch_adj_pair_a<false, false, false>(w, g, b, st, io);
}

template <int SPB>
void launch_adj_u(const AcDev &v, dim3 g, dim3 b, hipStream_t st, const float *in, float *out, int t,
                  const float *imd, const float *imp, float *ga1, float *ga2, float *gs) {
  switch (v.unroll) {
    case 4: hipLaunchKernelGGL((ac_adj_step<SPB, 4>), g, b, 0, st, v, in, out, t, imd, imp, ga1, ga2, gs); break;
    case 2: hipLaunchKernelGGL((ac_adj_step<SPB, 2>), g, b, 0, st, v, in, out, t, imd, imp, ga1, ga2, gs); break;
    default: hipLaunchKernelGGL((ac_adj_step<SPB, 1>), g, b, 0, st, v, in, out, t, imd, imp, ga1, ga2, gs);
  }
}

void launch_adj(const AcDev &v, dim3 g, dim3 b, hipStream_t st, const float *in, float *out, int t,
                const float *imd, const float *imp, float *ga1, float *ga2, float *gs) {
  switch (v.spb) {
    case 8: launch_adj_u<8>(v, g, b, st, in, out, t, imd, imp, ga1, ga2, gs); break;
    case 4: launch_adj_u<4>(v, g, b, st, in, out, t, imd, imp, ga1, ga2, gs); break;
    case 2: launch_adj_u<2>(v, g, b, st, in, out, t, imd, imp, ga1, ga2, gs); break;
    default: launch_adj_u<1>(v, g, b, st, in, out, t, imd, imp, ga1, ga2, gs);
  }
}

}  // namespace

extern "C" {

int adfwi_acoustic_get_layout(const adfwi_acoustic_desc *d, adfwi_layout *out) {
  ADFWI_CHECK_ARG(out != nullptr, "null layout pointer");
  return layout_of(d, out);
}

int adfwi_acoustic_forward(const adfwi_acoustic_desc *d, adfwi_acoustic_fwd_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_acoustic_forward")) return rc;
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null forward args");
  const bool recs = a->rec_p || a->rec_u || a->rec_w;
  ADFWI_CHECK_ARG(!recs || (a->rec_p && a->rec_u && a->rec_w), "rec_p/rec_u/rec_w must be all set or all NULL");
  if ((rc = check_device_desc(d, recs))) return rc;
  ADFWI_CHECK_ARG(a->state_in && a->work[0] && a->work[1], "null state/work buffer");
  ADFWI_CHECK_ARG(a->work[0] != a->work[1], "work[0] and work[1] must be distinct buffers");
  ADFWI_CHECK_ARG(a->t0 >= 0 && a->nsteps >= 0 && a->t0 + a->nsteps <= d->nt,
                  "steps [%d, %d) outside [0, %d)", a->t0, a->t0 + a->nsteps, d->nt);
  ADFWI_CHECK_ARG(!a->ckpt || a->ckpt_every >= 1, "ckpt_every must be >= 1 with a checkpoint buffer");
  ADFWI_CHECK_ARG(!a->t_base || !a->ckpt, "checkpoints are not written with a device time base");
  AcDev v = make_dev(d, L);
  v.tbase = a->t_base;
  v.eu_from = a->energy_u_from;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = step_grid(L), block(NT);
  const long long ng = (long long)d->nshots * d->nrec;
  const dim3 ggrid((unsigned)((ng + 255) / 256)), gblock(256);
  // two steps per launch where the first step's state is not needed in HBM (it is not a
  // checkpoint) and the receiver samples of both steps can be written through the receiver-index
  // map
  const bool fuse = a->fuse_steps != 0 && L.shots_per_block <= 4 && (!recs || d->nrec == 0 || a->rcv_map != nullptr);
  const bool drop_ok = pair_drop_ok(L);
  const AcDev w = pair_dev(v);
  const dim3 pgrid((w.nstrips + WPB - 1) / WPB, L.ngroups);
  auto ckpt_slot = [&](int t) -> float * {   // checkpoint buffer receiving the state after step t
    if (a->ckpt && (t + 1) % a->ckpt_every == 0 && t + 1 < d->nt)
      return a->ckpt + (long long)((t + 1) / a->ckpt_every - 1) * L.state_floats;
    return nullptr;
  };
  const float *cur = a->state_in;
  for (int k = 0; k < a->nsteps;) {
    const int t = a->t0 + k;
    const bool pair = fuse && k + 1 < a->nsteps && ckpt_slot(t) == nullptr;
    const int tl = pair ? t + 1 : t;         // last step of this launch
    float *nxt = ckpt_slot(tl);
    if (!nxt) nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    float *imd = a->image ? a->image + (long long)k * L.field_floats : nullptr;
    float *imp = a->image_p ? a->image_p + (long long)k * L.field_floats : nullptr;
    if (pair) {
      PairIO io;
      io.in = cur; io.out = nxt; io.t = t;
      io.energy = a->energy;
      io.img_d0 = imd; io.img_d1 = imd ? imd + L.field_floats : nullptr;
      io.img_p0 = imp; io.img_p1 = imp ? imp + L.field_floats : nullptr;
      io.rmap = a->rcv_map;
      io.rec_p = recs ? a->rec_p : nullptr; io.rec_u = a->rec_u; io.rec_w = a->rec_w;
      launch_pair(w, pgrid, block, st, io, drop_ok);
      if ((rc = adfwi::check_launch("ac_fwd_pair"))) return rc;
    } else {
      const int acc_u = a->energy ? 1 : 0;   // from step energy_u_from on (tested in the kernel)
      launch_fwd(v, grid, block, st, cur, nxt, t, a->energy, acc_u, imd, imp);
      if ((rc = adfwi::check_launch("ac_fwd_step"))) return rc;
    }
    if (recs && ng > 0 && !pair) {   // (a pair launch takes the records of both its steps)
      hipLaunchKernelGGL(ac_gather, ggrid, gblock, 0, st, v, nxt, a->rec_p, a->rec_u, a->rec_w, tl);
      if ((rc = adfwi::check_launch("ac_gather"))) return rc;
    }
    cur = nxt;
    k += pair ? 2 : 1;
  }
  a->state_out = cur;
  return ADFWI_OK;
}

int adfwi_acoustic_adjoint(const adfwi_acoustic_desc *d, adfwi_acoustic_adj_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_acoustic_adjoint")) return rc;
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null adjoint args");
  const bool recs = a->grad_rec_p || a->grad_rec_u || a->grad_rec_w;
  if ((rc = check_device_desc(d, recs))) return rc;
  ADFWI_CHECK_ARG(a->lam_in && a->work[0] && a->work[1], "null adjoint state/work buffer");
  ADFWI_CHECK_ARG(a->work[0] != a->work[1], "work[0] and work[1] must be distinct buffers");
  ADFWI_CHECK_ARG(a->lam_in == a->work[0] || a->lam_in == a->work[1],
                  "lam_in must be work[0] or work[1] (receiver sources are added in place)");
  ADFWI_CHECK_ARG(a->t_hi < d->nt && a->nsteps >= 0 && a->t_hi - a->nsteps + 1 >= 0,
                  "adjoint steps %d..%d outside [0, %d)", a->t_hi - a->nsteps + 1, a->t_hi, d->nt);
  ADFWI_CHECK_ARG(!a->grad_a1 || a->image, "grad_a1 needs the stencil image");
  ADFWI_CHECK_ARG(!a->grad_a2 || a->image_p, "grad_a2 needs the pressure image");
  ADFWI_CHECK_ARG(a->image_t0 <= a->t_hi - a->nsteps + 1 || !(a->image || a->image_p),
                  "image_t0 %d after the first adjoint step %d", a->image_t0, a->t_hi - a->nsteps + 1);
  AcDev v = make_dev(d, L);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = step_grid(L), block(NT);
  const long long ng = (long long)d->nshots * d->nrec;
  const dim3 ggrid((unsigned)((ng + 255) / 256)), gblock(256);
  // two steps per launch: the later one (t) reversed first, the receiver sources of t and t-1
  // added inside the kernel through the receiver-index map
  const bool fuse = a->fuse_steps != 0 && L.shots_per_block <= 4 && (!recs || d->nrec == 0 || a->rcv_map != nullptr);
  const bool drop_ok = pair_drop_ok(L);
  AcDev w = pair_dev(v);
  w.spb = L.reserved;
  const dim3 pgrid((w.nstrips + WPB - 1) / WPB, (d->nshots + w.spb - 1) / w.spb);
  // the one-step kernel runs at the finer of the two shot groupings, so that both adjoint
  // kernels add a shot's image terms into the same accumulator plane (plane = shot / finer):
  // the planes of a group then hold that group's shots only, whichever kernel reversed a step
  AcDev v1 = v;
  v1.spb = L.shots_per_block < L.reserved ? L.shots_per_block : L.reserved;
  const dim3 grid1(grid.x, (d->nshots + v1.spb - 1) / v1.spb);
  auto slot = [&](int t) -> long long { return (long long)(t - a->image_t0) * L.field_floats; };
  float *cur = const_cast<float *>(a->lam_in);
  for (int k = 0; k < a->nsteps;) {
    const int t = a->t_hi - k;
    const bool pair = fuse && k + 1 < a->nsteps;
    float *nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    AdjPairIO io{};
    if (pair) {
      io.in = cur; io.out = nxt; io.t = t - 1;
      io.img_d0 = a->image ? a->image + slot(t - 1) : nullptr;
      io.img_d1 = a->image ? a->image + slot(t) : nullptr;
      io.img_p0 = a->image_p ? a->image_p + slot(t - 1) : nullptr;
      io.img_p1 = a->image_p ? a->image_p + slot(t) : nullptr;
      io.grad_a1 = a->grad_a1; io.grad_a2 = a->grad_a2; io.grad_src = a->grad_src;
      io.rmap = recs ? a->rcv_map : nullptr;
      io.g_p = a->grad_rec_p; io.g_u = a->grad_rec_u; io.g_w = a->grad_rec_w;
    }
    // the record adjoints of step t: added in place before the launch, unless a pair kernel
    // adds them as it reads the state
    if (recs && ng > 0 && !(pair && adj_pair_adds_first(io))) {
      hipLaunchKernelGGL(ac_inject, ggrid, gblock, 0, st, v, cur, a->grad_rec_p, a->grad_rec_u, a->grad_rec_w, t);
      if ((rc = adfwi::check_launch("ac_inject"))) return rc;
    }
    if (pair) {
      launch_adj_pair(w, pgrid, block, st, io, drop_ok);
      if ((rc = adfwi::check_launch("ac_adj_pair"))) return rc;
    } else {
      const float *imd = a->image ? a->image + slot(t) : nullptr;
      const float *imp = a->image_p ? a->image_p + slot(t) : nullptr;
      launch_adj(v1, grid1, block, st, cur, nxt, t, imd, imp, a->grad_a1, a->grad_a2, a->grad_src);
      if ((rc = adfwi::check_launch("ac_adj_step"))) return rc;
    }
    cur = nxt;
    k += pair ? 2 : 1;
  }
  a->lam_out = cur;
  return ADFWI_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free adjoint (adfwi_acoustic_adjoint_noalloc)
// ---------------------------------------------------------------------------
// The adjoint kernels add each shot group's zero-lag image terms into a plane of its own
// (deterministic, no atomics); the caller of adfwi_acoustic_adjoint supplies those planes.
// Callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc) use the variant below instead: the planes live in this
// file-scope array, are cleared by a fill kernel, and are summed in a fixed order into the
// caller's one-plane outputs after the time steps.  Each GPU holds its own copy of the array
// until the process ends; every call on a GPU shares it, so calls that overlap on one GPU
// (two streams, two threads) race on it.
#include <mutex>

namespace {

constexpr long long NOALLOC_SCRATCH_FLOATS = ADFWI_NOALLOC_SCRATCH_FLOATS;   // 128 MiB per GPU
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];

std::mutex g_scratch_mu;
float *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.  The host
// symbol `scratch` is only a stand-in and must never be passed to a kernel.
int scratch_of_current_device(float **out) {
  int dev = -1;
  ADFWI_CHECK_HIP(hipGetDevice(&dev));
  ADFWI_CHECK_ARG(dev >= 0 && dev < MAX_DEVICES, "device %d outside [0, %d)", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    ADFWI_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ADFWI_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = static_cast<float *>(p);
  }
  *out = g_scratch_ptr[dev];
  return ADFWI_OK;
}

__global__ void fill_zero(float *__restrict__ p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// out[i] += planes[0][i] + planes[1][i] + ... in plane order (deterministic)
__global__ void add_plane_sum(const float *__restrict__ planes, int nplanes, long long plane, float *__restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < plane; i += (long long)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int g = 0; g < nplanes; ++g) s += planes[g * plane + i];
    out[i] += s;
  }
}

dim3 flat_grid(long long n) {
  long long b = (n + 255) / 256;
  return dim3((unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096));
}

}  // namespace

extern "C" {

int adfwi_acoustic_adjoint_noalloc(const adfwi_acoustic_desc *d, adfwi_acoustic_adj_args *a, float *grad_a1_sum,
                                   float *grad_a2_sum, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_acoustic_adjoint_noalloc")) return rc;
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null adjoint args");
  ADFWI_CHECK_ARG(!a->grad_a1 && !a->grad_a2,
                  "grad_a1/grad_a2 must be NULL: the per-group planes live in the library's scratch; "
                  "pass grad_a1_sum/grad_a2_sum");
  ADFWI_CHECK_ARG(!grad_a1_sum || a->image, "grad_a1_sum needs the stencil image");
  ADFWI_CHECK_ARG(!grad_a2_sum || a->image_p, "grad_a2_sum needs the pressure image");
  const long long plane = (long long)d->nz * d->pitch;
  const long long nplanes = L.accum_floats / plane;
  const long long need = nplanes * plane * ((grad_a1_sum ? 1 : 0) + (grad_a2_sum ? 1 : 0));
  ADFWI_CHECK_ARG(need <= NOALLOC_SCRATCH_FLOATS,
                  "gradient planes need %lld floats of scratch (%lld planes of %d x %d), the maximum is %lld "
                  "(128 MiB); use adfwi_acoustic_adjoint with caller-owned planes",
                  need, nplanes * ((grad_a1_sum ? 1 : 0) + (grad_a2_sum ? 1 : 0)), d->nz, d->pitch,
                  NOALLOC_SCRATCH_FLOATS);
  float *base = nullptr;
  if (need > 0 && (rc = scratch_of_current_device(&base))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  adfwi_acoustic_adj_args b = *a;
  b.grad_a1 = grad_a1_sum ? base : nullptr;
  b.grad_a2 = grad_a2_sum ? base + (grad_a1_sum ? nplanes * plane : 0) : nullptr;
  if (need > 0) {
    hipLaunchKernelGGL(fill_zero, flat_grid(need), dim3(256), 0, st, base, need);
    if ((rc = adfwi::check_launch("fill_zero"))) return rc;
  }
  if ((rc = adfwi_acoustic_adjoint(d, &b, stream))) return rc;
  a->lam_out = b.lam_out;
  if (grad_a1_sum) {
    hipLaunchKernelGGL(add_plane_sum, flat_grid(plane), dim3(256), 0, st, b.grad_a1, (int)nplanes, plane, grad_a1_sum);
    if ((rc = adfwi::check_launch("add_plane_sum"))) return rc;
  }
  if (grad_a2_sum) {
    hipLaunchKernelGGL(add_plane_sum, flat_grid(plane), dim3(256), 0, st, b.grad_a2, (int)nplanes, plane, grad_a2_sum);
    if ((rc = adfwi::check_launch("add_plane_sum"))) return rc;
  }
  return ADFWI_OK;
}

}  // extern "C"
