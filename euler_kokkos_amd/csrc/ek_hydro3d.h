// ek_hydro3d.h -- fused 3D hydro MUSCL-Hancock step for gfx950, marching along z.
//
// One block owns a TX x TY column tile and a chunk of z planes. Every thread owns one
// position of the tile's (TX+4) x (TY+4) primitive grid for the whole march and keeps
// that column's primitives at planes p-1, p, p+1 in registers, so the z slope never
// touches LDS; only plane p (for the x/y slopes) and the face states live in LDS
// (about 42 KB for 16x8 in FP64, three blocks per CU).
//
// Per plane p:  A  load U(p+1), primitives into registers (all positions)
//               B  slopes + Hancock trace at the tile and its 1-wide halo cross;
//                  x/y face states to LDS; each tile cell parks its two z-face states
//                  (max-z of plane p-1, min-z of plane p) in its own LDS slot
//               -- barrier --
//               W  primitives of plane p+1 into the LDS plane
//               C  Riemann problems: every tile cell solves its own z face (in place, its
//                  slot only); the (TX+1)TY x-faces and TX(TY+1) y-faces are split over the
//                  waves so that each wave's share of this phase is even; x/y fluxes go to
//                  their own buffer
//               -- barrier --
//               D  update of plane p (UpdateFunctor3D order) and the -Fz(p) term plus the
//                  CFL contribution of plane p-1
// D reads only the x/y flux buffer (rewritten after the next plane's first barrier) and
// the cell's own z slot, so no third barrier is needed: D of plane p overlaps A/B of p+1.
// The arithmetic per cell, face and update is exactly the reference's
// (src/muscl/HydroRunFunctors3D.h:271-820); only where and how often it is evaluated
// changes: slopes once per cell (the reference recomputes them for 4 cells), each face
// solved once.
#pragma once

#include "ek_hydro.h"

namespace ek {

// Column tile of the single-grid 3D step: 16 x 8 cells, 256 threads.
constexpr int T3X = 16, T3Y = 8, N3T = 256;

// z-chunk length of the single-grid 3D step for an nx x ny x nz grid: the longest of 32,
// 24, 16, 12, 8 that still gives 2048 blocks (8 per CU, enough to fill the 256 CUs with a
// short tail), else 4. Each chunk re-traces two planes it does not update, so the longest
// chunk that fills the GPU is fastest; on small grids shorter chunks are what fills it
// (64^3: 0.134 ms per step at 32, 0.037 at 4; 192^3: 0.551 at 32, 0.532 at 24; 256^3 keeps 32).
// The planes are then spread evenly over that many chunks: a step lasts as long as its
// longest chunk (256^3: 1.20 ms at 32 planes, 1.25 at 30 or 34 and 1.33 at 36, whose last
// chunk is short).
inline int hydro3d_z_chunk(int nx, int ny, int nz)
{
  const long long tiles = (long long)((nx + T3X - 1) / T3X) * ((ny + T3Y - 1) / T3Y);
  int cz = 4;
  for (const int c : { 32, 24, 16, 12, 8 })
    if (tiles * ((nz + c - 1) / c) >= 2048) {
      cz = c;
      break;
    }
  if (cz >= nz)
    return nz;
  const int nch = (nz + cz - 1) / cz;
  return (nz + nch - 1) / nch;
}

template <int TX, int TY, int NT>
struct Col3 {
  static constexpr int QX = TX + 4, QY = TY + 4, QN = QX * QY;
  static constexpr int PT = TX * TY;               // tile cells (all six face states)
  static constexpr int PH = 2 * TX + 2 * TY;       // halo cross (one face state each)
  static constexpr int PO = 2 * TX + 2 * TY + 4;   // primitives only (slope stencils)
  static constexpr int NPOS = PT + PH + PO;
  static constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1);
  static constexpr int NXY = NFX + NFY;            // x and y faces (shared list)
  static constexpr int NF = NXY + PT;              // plus one z slot per tile cell
  static constexpr int NW = NT / 64;
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(NPOS <= NT, "one primitive column per thread");
};

// Static split of the x/y face list over the waves. Tile-cell waves already solve their
// own z faces in the Riemann phase (one round); each further 64-face round goes to the
// wave with the fewest rounds in that phase, ties to the one with less work before it
// (trace, update), so the phase between the two barriers is as short as it can be.
template <int TX, int TY, int NT>
struct FacePlan {
  int start[NT / 64];
  int count[NT / 64];
};

template <int TX, int TY, int NT>
constexpr FacePlan<TX, TY, NT> make_face_plan()
{
  using C = Col3<TX, TY, NT>;
  FacePlan<TX, TY, NT> fp{};
  int before[C::NW] = {};  // work before the first barrier: update of p-1, trace of p
  int rounds[C::NW] = {};  // 64-face rounds between the barriers
  int xy[C::NW] = {};
  for (int w = 0; w < C::NW; w++) {
    const int lo = 64 * w, hi = 64 * w + 64;
    const bool tile = lo < C::PT;
    const bool cross = hi > C::PT && lo < C::PT + C::PH;
    before[w] = tile ? 34 + 15 : (cross ? 29 : 0);
    rounds[w] = tile ? 1 : 0;
  }
  const int R = (C::NXY + 63) / 64;
  for (int r = 0; r < R; r++) {
    int best = 0;
    for (int w = 1; w < C::NW; w++)
      if (rounds[w] < rounds[best] || (rounds[w] == rounds[best] && before[w] < before[best]))
        best = w;
    rounds[best] += 1;
    xy[best] += 1;
  }
  // contiguous ranges; the wave with most work before the barrier goes last and so takes
  // the partial round
  int order[C::NW] = {};
  for (int w = 0; w < C::NW; w++)
    order[w] = w;
  for (int i = 0; i < C::NW; i++)
    for (int j = i + 1; j < C::NW; j++)
      if (before[order[j]] < before[order[i]]) {
        const int t = order[i];
        order[i] = order[j];
        order[j] = t;
      }
  int s = 0;
  for (int i = 0; i < C::NW; i++) {
    const int w = order[i];
    fp.start[w] = s;
    int c = xy[w] * 64;
    if (s + c > C::NXY)
      c = C::NXY - s;
    if (c < 0)
      c = 0;
    fp.count[w] = c;
    s += c;
  }
  return fp;
}

// Q-grid coordinates of position t and the face states it must produce
// (bit 2d: min face along d, bit 2d+1: max face along d).
template <int TX, int TY, int NT>
__device__ __forceinline__ bool col_pos(int t, int& qx, int& qy, unsigned& want)
{
  using C = Col3<TX, TY, NT>;
  if (t < C::PT) {
    qx = t % TX + 2;
    qy = t / TX + 2;
    want = 0x3F;
    return true;
  }
  if (t < C::PT + C::PH) {
    const int h = t - C::PT;
    if (h < TY) {
      qx = 1; qy = h + 2; want = 0x2;
    } else if (h < 2 * TY) {
      qx = TX + 2; qy = h - TY + 2; want = 0x1;
    } else if (h < 2 * TY + TX) {
      qx = h - 2 * TY + 2; qy = 1; want = 0x8;
    } else {
      qx = h - 2 * TY - TX + 2; qy = TY + 2; want = 0x4;
    }
    return true;
  }
  if (t < C::NPOS) {
    const int o = t - C::PT - C::PH;
    want = 0;
    if (o < TY) {
      qx = 0; qy = o + 2;
    } else if (o < 2 * TY) {
      qx = TX + 3; qy = o - TY + 2;
    } else if (o < 2 * TY + TX) {
      qx = o - 2 * TY + 2; qy = 0;
    } else if (o < 2 * TY + 2 * TX) {
      qx = o - 2 * TY - TX + 2; qy = TY + 3;
    } else {
      const int c = o - 2 * TY - 2 * TX;
      qx = (c & 1) ? TX + 2 : 1;
      qy = (c & 2) ? TY + 2 : 1;
    }
    return true;
  }
  qx = qy = 0;
  want = 0;
  return false;
}

// Component permutations of the y and z solves: IU <-> IV, IU <-> IW (self-inverse).
__device__ __forceinline__ constexpr int rot_y(int v)
{
  return v == IU ? IV : (v == IV ? IU : v);
}
__device__ __forceinline__ constexpr int rot_z(int v)
{
  return v == IU ? IW : (v == IW ? IU : v);
}

// One TX x TY column tile over planes [k0, k0+CZ) of the 3D step with time step dt.
// Returns this thread's contribution to the CFL maximum of the new state. Shared by the
// single-grid kernel below and the batched (job pool) kernel. EARLY: the owner reads its
// U(p) before the Riemann phase instead of in the update (below).
template <class T, int TX, int TY, int NT, int RS, int ST, bool EARLY = (sizeof(T) == 8)>
__device__ __forceinline__ double hydro3d_column(const T* __restrict__ Uin, T* __restrict__ Uout,
                                                 const Geom<T>& geo, const Gas<T>& gas, double dt, int i0,
                                                 int j0, int k0, int CZ)
{
  using C = Col3<TX, TY, NT>;
  constexpr int NV = 5;
  constexpr int QX = C::QX, QN = C::QN, PT = C::PT;
  constexpr int NFX = C::NFX, NXY = C::NXY, NF = C::NF;
  constexpr FacePlan<TX, TY, NT> plan = make_face_plan<TX, TY, NT>();
  __shared__ T sQ[NV][QN];
  // Face states, rotated so that IU holds the normal velocity (the swap the reference does
  // before each y/z solve, HydroRunFunctors3D.h:597,692): [0] left state, [1] right state.
  // Faces: x (fy*(TX+1)+fx), then y (NFX + fy*TX+fx), then z (NXY + tile cell). A z slot
  // is touched only by its own tile cell's thread; after its solve, [0][.][z] holds the
  // z flux times dt/dz, still rotated.
  __shared__ T sF[2][NV][NF];
  // x/y fluxes times dt/dx (dt/dy), rotated like the states
  __shared__ T sG[NV][NXY];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const T dtdx[3] = { T(dt / geo.dx), T(dt / geo.dy), T(dt / geo.dz) };
  const int nx = geo.nx, ny = geo.ny, nz = geo.nz;
  const int k1 = min(k0 + CZ, nz);

  int qx, qy;
  unsigned want;
  const bool haspos = col_pos<TX, TY, NT>(tid, qx, qy, want);
  const int c = qy * QX + qx;
  bool negx, negy;
  const int si = bc_map(i0 - 2 + qx, nx, geo.bc[0], geo.bc[1], negx);
  const int sj = bc_map(j0 - 2 + qy, ny, geo.bc[2], geo.bc[3], negy);
  const long long col = (long long)sj * nx + si;
  const bool owner = tid < PT;
  const int ox = tid % TX, oy = tid / TX;
  const bool ovalid = owner && (i0 + ox) < nx && (j0 + oy) < ny;
  const long long ocol = (long long)(j0 + oy) * nx + (i0 + ox);

  T qp[NV], qc[NV], qn[NV];  // this column's primitives at planes p-1, p, p+1
  T qzm[NV];                 // owner: max-z face state of its cell in plane p-1
  T acc[NV];                 // owner: its cell in plane p-1, updated up to +Fz(p-1)
  double local = 0.0;

  // raw conservative state of this column at plane kp, ghost planes through the boundary map
  auto fetch = [&](int kp, T* u) {
    bool negz;
    const int sk = bc_map(kp, nz, geo.bc[4], geo.bc[5], negz);
    const long long o = (long long)sk * geo.plane + col;
#pragma unroll
    for (int v = 0; v < NV; v++)
      u[v] = Uin[v * geo.vstride + o];
    if (negx)
      u[IU] = u[IU] * T(-1);
    if (negy)
      u[IV] = u[IV] * T(-1);
    if (negz)
      u[IW] = u[IW] * T(-1);
  };

  // software pipeline: upre holds U(p+1) when iteration p starts; the loads for p+2 are
  // issued right after it is consumed and complete behind phases B-D
  T upre[NV], ucur[NV];
  if (haspos) {
    fetch(k0 - 2, upre);
    prims<T, 3>(gas, upre, qp);
    fetch(k0 - 1, upre);
    prims<T, 3>(gas, upre, qc);
#pragma unroll
    for (int v = 0; v < NV; v++)
      sQ[v][c] = qc[v];
    fetch(k0, upre);
  }
  __syncthreads();

  for (int p = k0 - 1; p <= k1; p++) {
    const bool warm = p == k0 - 1, last = p == k1;
    // A: primitives of plane p+1 for this column; prefetch p+2
    if (haspos) {
      prims<T, 3>(gas, upre, qn);
      if (p + 2 <= k1 + 1)
        fetch(p + 2, upre);
    }
    // B: slopes and face states of plane p. The two tile-cell waves trace six faces per
    // cell, the halo waves one or none and then wait at the barrier: the tile-cell waves
    // take issue priority until it (FP64 256^3: 1.190 -> 1.175 ms per step)
    const bool heavy = __builtin_amdgcn_readfirstlane(wave) < 2;
    if (heavy)
      __builtin_amdgcn_s_setprio(2);
    if (want) {
      unsigned w = want;
      if (warm || last)
        w = owner ? (warm ? 0x20u : 0x10u) : 0u;
      if (w) {
        T dq[3][5];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          if constexpr (ST == 0) {
            dq[0][v] = T(0);
            dq[1][v] = T(0);
            dq[2][v] = T(0);
          } else {
            dq[2][v] = slope1<T, ST>(qc[v], qn[v], qp[v]);
            dq[0][v] = slope1<T, ST>(qc[v], sQ[v][c + 1], sQ[v][c - 1]);
            dq[1][v] = slope1<T, ST>(qc[v], sQ[v][c + QX], sQ[v][c - QX]);
          }
        }
        T s[NV];
        trace_source<T, 3>(gas, qc, dq, s);
        const int cx = qx - 2, cy = qy - 2;
        // rotated stores: component IU <- normal velocity of the face direction
        if (w & 0x2) {
          T f[NV];
          trace_face<T, NV>(gas, qc, dq[0], s, dtdx[0], true, f);
          const int fi = cy * (TX + 1) + cx + 1;
#pragma unroll
          for (int v = 0; v < NV; v++)
            sF[0][v][fi] = f[v];
        }
        if (w & 0x1) {
          T f[NV];
          trace_face<T, NV>(gas, qc, dq[0], s, dtdx[0], false, f);
          const int fi = cy * (TX + 1) + cx;
#pragma unroll
          for (int v = 0; v < NV; v++)
            sF[1][v][fi] = f[v];
        }
        if (w & 0x8) {
          T f[NV];
          trace_face<T, NV>(gas, qc, dq[1], s, dtdx[1], true, f);
          const int fi = NFX + (cy + 1) * TX + cx;
#pragma unroll
          for (int v = 0; v < NV; v++)
            sF[0][v][fi] = f[rot_y(v)];
        }
        if (w & 0x4) {
          T f[NV];
          trace_face<T, NV>(gas, qc, dq[1], s, dtdx[1], false, f);
          const int fi = NFX + cy * TX + cx;
#pragma unroll
          for (int v = 0; v < NV; v++)
            sF[1][v][fi] = f[rot_y(v)];
        }
        if (w & 0x10) {
          T f[NV];
          trace_face<T, NV>(gas, qc, dq[2], s, dtdx[2], false, f);
          const int fi = NXY + tid;
#pragma unroll
          for (int v = 0; v < NV; v++) {
            sF[0][v][fi] = qzm[rot_z(v)];
            sF[1][v][fi] = f[rot_z(v)];
          }
        }
        if (w & 0x20)
          trace_face<T, NV>(gas, qc, dq[2], s, dtdx[2], true, qzm);
      }
    }
    if (heavy)
      __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    // W: plane p+1 becomes the LDS plane (every read of plane p is done)
    if (haspos && !last) {
#pragma unroll
      for (int v = 0; v < NV; v++)
        sQ[v][c] = qn[v];
    }
    // EARLY (the FP64 single-grid step): the owner's U(p) is read here, so the load
    // completes behind the Riemann phase; the 10 registers it holds through C stay under the
    // 3-wave limit that the LDS footprint sets anyway (in FP32 they would cost a wave, and
    // the pool kernel, with its job records, would spill)
    if constexpr (EARLY) {
      if (owner && !warm && !last && ovalid) {
        const long long o = (long long)p * geo.plane + ocol;
#pragma unroll
        for (int v = 0; v < NV; v++)
          ucur[v] = Uin[v * geo.vstride + o];
      }
    }
    // C: Riemann problems of plane p's faces: own z face, then this wave's x/y range
    if (!warm) {
      if (owner) {
        const int f = NXY + tid;
        T ql[NV], qr[NV], fl[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          ql[v] = sF[0][v][f];
          qr[v] = sF[1][v][f];
        }
        riemann<T, NV, RS>(gas, ql, qr, fl);
#pragma unroll
        for (int v = 0; v < NV; v++)
          sF[0][v][f] = fl[v] * dtdx[2];
      }
      if (!last) {
        const int fe = plan.start[wave] + plan.count[wave];
        for (int f = plan.start[wave] + lane; f < fe; f += 64) {
          T ql[NV], qr[NV], fl[NV];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            ql[v] = sF[0][v][f];
            qr[v] = sF[1][v][f];
          }
          riemann<T, NV, RS>(gas, ql, qr, fl);
          const T dtd = f < NFX ? dtdx[0] : dtdx[1];
#pragma unroll
          for (int v = 0; v < NV; v++)
            sG[v][f] = fl[v] * dtd;
        }
      }
    }
    __syncthreads();
    // D: finish plane p-1 (-Fz(p)), accumulate plane p (+Fx(i) -Fx(i+1) +Fy(j) -Fy(j+1) +Fz(p))
    if (owner && !warm) {
      if (p > k0) {
        T u[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          acc[v] -= sF[0][rot_z(v)][NXY + tid];
          u[v] = acc[v];
        }
        if (ovalid) {
          const long long o = (long long)(p - 1) * geo.plane + ocol;
#pragma unroll
          for (int v = 0; v < NV; v++)
            Uout[v * geo.vstride + o] = u[v];
          const double s = cfl_term<T, 3>(gas, u, geo);
          local = (s > local || !(s == s)) ? s : local;
        }
      }
      if (!last) {
        // the owner's U(p), read late to keep it out of the register peak of B and C (this
        // thread streamed the same column two planes ago, so the read hits in cache)
        if constexpr (!EARLY) {
          if (ovalid) {
            const long long o = (long long)p * geo.plane + ocol;
#pragma unroll
            for (int v = 0; v < NV; v++)
              ucur[v] = Uin[v * geo.vstride + o];
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
          T a = ovalid ? ucur[v] : T(0);
          a += sG[v][oy * (TX + 1) + ox];
          a -= sG[v][oy * (TX + 1) + ox + 1];
          a += sG[rot_y(v)][NFX + oy * TX + ox];
          a -= sG[rot_y(v)][NFX + (oy + 1) * TX + ox];
          a += sF[0][rot_z(v)][NXY + tid];
          acc[v] = a;
        }
      }
    }
    // no barrier here: D reads sG (next written after the next plane's first barrier) and
    // this thread's own z slot
#pragma unroll
    for (int v = 0; v < NV; v++) {
      qp[v] = qc[v];
      qc[v] = qn[v];
    }
  }
  return local;
}

template <class T, int TX, int TY, int NT, int RS, int ST>
__global__ __launch_bounds__(NT) void hydro3d_step(const T* __restrict__ Uin, T* __restrict__ Uout,
                                                   Geom<T> geo, Gas<T> gas, StepCtl ctl, int CZ,
                                                   int ntx, int nty, int nchunk)
{
  __shared__ double sred[NT / 64];
  // XCD-aware tile order: blocks are dispatched round-robin over the 8 XCDs, so give each
  // XCD a contiguous range of tiles (neighbours share halo lines in that XCD's L2).
  const int nblk = ntx * nty * nchunk;
  const int b = blockIdx.x;
  const int per = nblk / 8;
  const int lb = (per * 8 == nblk) ? (b % 8) * per + b / 8 : b;
  const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
  const bool first_block = b == 0;
  double dt;
  TimeSlot cur;
  if (!step_begin(ctl, first_block, dt, cur))
    return;
  const double local =
      hydro3d_column<T, TX, TY, NT, RS, ST>(Uin, Uout, geo, gas, dt, bx * TX, by * TY, bz * CZ, CZ);
  block_max_to<NT>(invdt_word(ctl, (ctl.slot + 1) % 3, b), local, sred);
  if (first_block && threadIdx.x < 64)
    step_end_bookkeeping(ctl, cur, dt);
}

} // namespace ek
