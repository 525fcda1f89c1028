// grad_prim.h -- derivative-ERI contraction with one 64-lane wave per
// contracted quartet and one lane per primitive quartet.
//
// The G block of the quartet and the Hermite E tables of its bra and ket
// primitive pairs sit in LDS (each table built once per primitive pair by
// three lanes); every lane then takes one primitive quartet (bra primitive
// ib, ket primitive ik) through the whole MMD derivative contraction on its
// own: Boys values and R_tuv up to order L+1 in registers, and per ket
// Cartesian pair cd the ket-contracted Hermite integrals
//   h0_tuv = sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}
// (t+u+v <= la+lb+1) and the C-derivative ones hC_{x,y,z} (t+u+v <= la+lb),
// which each bra Cartesian pair ab takes against its A-derivative,
// P-shifted and plain E products, weighted by G_abcd (read from LDS, the
// same address for every lane).  The lanes' nine sums (dA, dP = d(A+B),
// dC) are reduced over the wave at the end.  Compared with the one-wave-
// per-quartet kernel (grad_impl.h), whose lanes share the work of one
// primitive quartet at a time between barriers, there is no barrier inside
// the primitive loop: the per-quartet serial phases disappear, at the price
// of R and the h lists held per lane.
#pragma once

#include <type_traits>

#include "grad_impl.h"

namespace qcmd {

template <int LA, int LB, int LC, int LD, typename Real>
struct GradPrim {
  static constexpr int LAB = LA + LB, LCD = LC + LD, L1 = LAB + LCD + 1;
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NAB = NA * NB, NCD = NC * ND;
  static constexpr int NH1 = nherm(LAB + 1), NH0 = nherm(LAB);
  static constexpr int EBS = (LA + 2) * (LB + 1) * (LAB + 2);  // E^{ij}_t: i <= la+1, j <= lb
  static constexpr int EKS = (LC + 2) * (LD + 1) * (LCD + 2);
  struct Shared {
    Real gam[NAB * NCD];
    Real eb[GMAXP][3][EBS];
    Real ek[GMAXP][3][EKS];
  };
};

// Compile-time loops: f(std::integral_constant<int, i>) for i in [B, E), so
// that every index derived from i is a constant expression and the arrays
// it indexes stay in registers however large the unrolled nest is.
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

struct TUV {
  int t, u, v;
};
// (t, u, v) of entry k of the degree-ordered tetrahedron (herm_index inverse)
__host__ __device__ constexpr TUV herm_decode(int k) {
  for (int s = 0; s < 32; ++s)
    for (int t = s; t >= 0; --t)
      for (int u = s - t; u >= 0; --u)
        if (herm_index(t, u, s - t - u) == k) return TUV{t, u, s - t - u};
  return TUV{0, 0, 0};
}

// R^0_tuv (t+u+v <= L) from F_n by the MMD level recursion (the recurrences
// of hermite_r in qcmd_device.h), in place over the tetrahedron: level n
// rewrites the entries of degree L-n down to 1 in descending order, so each
// reads the level-(n+1) values of lower degree before they are replaced.
template <int L, typename Real>
__device__ __forceinline__ void hermite_r_tet(Real (&R)[nherm(L)], const Real (&F)[L + 1], Real m2a,
                                              Real X, Real Y, Real Z) {
  Real pw[L + 1];
  pw[0] = (Real)1;
#pragma unroll
  for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
  R[0] = pw[L] * F[L];
  sfor<0, L>([&](auto nI) {
    constexpr int n = L - 1 - decltype(nI)::value;
    sfor<0, nherm(L - n) - 1>([&](auto jI) {
      constexpr int k = nherm(L - n) - 1 - decltype(jI)::value;
      constexpr TUV e = herm_decode(k);
      Real r;
      if constexpr (e.t > 0) {
        r = X * R[herm_index(e.t - 1, e.u, e.v)];
        if constexpr (e.t > 1) r += (Real)(e.t - 1) * R[herm_index(e.t - 2, e.u, e.v)];
      } else if constexpr (e.u > 0) {
        r = Y * R[herm_index(e.t, e.u - 1, e.v)];
        if constexpr (e.u > 1) r += (Real)(e.u - 1) * R[herm_index(e.t, e.u - 2, e.v)];
      } else {
        r = Z * R[herm_index(e.t, e.u, e.v - 1)];
        if constexpr (e.v > 1) r += (Real)(e.v - 1) * R[herm_index(e.t, e.u, e.v - 2)];
      }
      R[k] = r;
    });
    R[0] = pw[n] * F[n];
  });
}

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CD1>
__global__ __launch_bounds__(64) void grad2e_prim_kernel(Grad2eArgs args) {
  using W = GradPrim<LA, LB, LC, LD, Real>;
  constexpr int LAB = W::LAB, LCD = W::LCD, L1 = W::L1;
  constexpr int NA = W::NA, NB = W::NB, NC = W::NC, ND = W::ND;
  constexpr int NAB = W::NAB, NCD = W::NCD, NH1 = W::NH1, NH0 = W::NH0;
  __shared__ typename W::Shared sh;

  const int lane = threadIdx.x;
  const int64_t gid = flat_block();
  if (gid >= args.nquartets) return;  // whole work-group (one wave)
  int lo = 0, hi = args.nbra - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.offsets[mid] <= gid)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int bp = args.bra_list[lo];
  const int kp = args.ket_list[args.kstart[lo] + (int)(gid - args.offsets[lo])];
  const Real* __restrict__ pp = gpairs<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];
  const int64_t oa = args.sp_aoc[2 * bp], ob = args.sp_aoc[2 * bp + 1];
  const int64_t oc = args.sp_aoc[2 * kp], od = args.sp_aoc[2 * kp + 1];
  const int64_t nc = args.ncart;

  // G block (and its max for the density screen, decided for the whole wave)
  double gmax = 0.0;
  for (int e = lane; e < NAB * NCD; e += 64) {
    const int ab = e / NCD, cd = e % NCD;
    const int64_t mu = oa + ab / NB, nu = ob + ab % NB, la = oc + cd / ND, si = od + cd % ND;
    double k = args.dk0[mu * nc + la] * args.dk0[nu * nc + si] +
               args.dk0[mu * nc + si] * args.dk0[nu * nc + la];
    if (args.nk > 1)
      k += args.dk1[mu * nc + la] * args.dk1[nu * nc + si] +
           args.dk1[mu * nc + si] * args.dk1[nu * nc + la];
    const Real gv = (Real)(args.dj[mu * nc + nu] * args.dj[la * nc + si] - args.cx * k);
    sh.gam[e] = gv;
    gmax = fmax(gmax, fabs((double)gv));
  }
  if (args.tau_density > 0.0) {
    for (int w = 32; w > 0; w >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, w));
    if (gmax * args.schwarz[bp] * args.schwarz[kp] < args.tau_density) return;  // uniform
  }

  Real acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = (Real)0;

  for (int b0 = 0; b0 < npb; b0 += GMAXP) {
    const int nb = min(GMAXP, npb - b0);
    for (int k0 = 0; k0 < npk; k0 += GMAXP) {
      const int nk = min(GMAXP, npk - k0);
      __syncthreads();  // the previous block's tables are no longer read
      if (lane < 3 * nb) {
        const int ib = lane / 3, dim = lane - 3 * ib;
        const Real* P = pp + (int64_t)(pb0 + b0 + ib) * PP_STRIDE;
        Real E[LA + 2][LB + 1][LAB + 2];
        hermite_e<LA + 1, LB, Real>(E, P[PP_PAX + dim], P[PP_PBX + dim], (Real)0.5 / P[PP_P]);
#pragma unroll
        for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LAB + 1; ++t)
              sh.eb[ib][dim][(i * (LB + 1) + j) * (LAB + 2) + t] = E[i][j][t];
      } else if (lane >= 32 && lane < 32 + 3 * nk) {
        const int ik = (lane - 32) / 3, dim = lane - 32 - 3 * ik;
        const Real* Q = pp + (int64_t)(pk0 + k0 + ik) * PP_STRIDE;
        Real E[LC + 2][LD + 1][LCD + 2];
        hermite_e<LC + 1, LD, Real>(E, Q[PP_PAX + dim], Q[PP_PBX + dim], (Real)0.5 / Q[PP_P]);
#pragma unroll
        for (int i = 0; i <= LC + 1; ++i)
#pragma unroll
          for (int j = 0; j <= LD; ++j)
#pragma unroll
            for (int t = 0; t <= LCD + 1; ++t)
              sh.ek[ik][dim][(i * (LD + 1) + j) * (LCD + 2) + t] = E[i][j][t];
      }
      __syncthreads();

      for (int q = lane; q < nb * nk; q += 64) {
        const int ib = q / nk, ik = q - ib * nk;
        const Real* P = pp + (int64_t)(pb0 + b0 + ib) * PP_STRIDE;
        const Real* Q = pp + (int64_t)(pk0 + k0 + ik) * PP_STRIDE;
        const Real p = P[PP_P], qq = Q[PP_P];
        const Real al2 = (Real)2 * P[PP_ALPHA], ga2 = (Real)2 * Q[PP_ALPHA];
        const Real pqm = p * qq, ps = p + qq, alpha = pqm / ps;
        const Real X = P[PP_PX] - Q[PP_PX], Y = P[PP_PX + 1] - Q[PP_PX + 1],
                   Z = P[PP_PX + 2] - Q[PP_PX + 2];
        Real F[L1 + 1];
        boys_eval<L1, Real>(alpha * (X * X + Y * Y + Z * Z), args.boys64, F);
        Real R[nherm(L1)];
        hermite_r_tet<L1, Real>(R, F, (Real)-2 * alpha, X, Y, Z);
        const Real pref = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * P[PP_K] * Q[PP_K];
        const Real* __restrict__ ebp = &sh.eb[ib][0][0];
        const Real* __restrict__ ekp = &sh.ek[ik][0][0];
        sfor<CD0, CD1>([&](auto cdI) {
          constexpr int c = decltype(cdI)::value / ND, d = decltype(cdI)::value % ND;
          constexpr int kx = cart_lx(LC, c), ky = cart_ly(LC, c), kz = cart_lz(LC, c);
          {
            constexpr int sx = kx + cart_lx(LD, d), sy = ky + cart_ly(LD, d), sz = kz + cart_lz(LD, d);
            constexpr int jx = cart_lx(LD, d), jy = cart_ly(LD, d), jz = cart_lz(LD, d);
            // this ket pair's E (and C-derivative E) factors per direction
            Real ex[sx + 2], ey[sy + 2], ez[sz + 2], dx[sx + 2], dy[sy + 2], dz[sz + 2];
            sfor<0, sx + 2>([&](auto tI) {
              constexpr int tt = decltype(tI)::value;
              ex[tt] = ekp[0 * W::EKS + (kx * (LD + 1) + jx) * (LCD + 2) + tt];
              {
                Real v = ga2 * ekp[0 * W::EKS + ((kx + 1) * (LD + 1) + jx) * (LCD + 2) + tt];
                if constexpr (kx > 0) v -= (Real)kx * ekp[0 * W::EKS + ((kx - 1) * (LD + 1) + jx) * (LCD + 2) + tt];
                dx[tt] = v;
              }
            });
            sfor<0, sy + 2>([&](auto uI) {
              constexpr int uu = decltype(uI)::value;
              ey[uu] = ekp[1 * W::EKS + (ky * (LD + 1) + jy) * (LCD + 2) + uu];
              {
                Real v = ga2 * ekp[1 * W::EKS + ((ky + 1) * (LD + 1) + jy) * (LCD + 2) + uu];
                if constexpr (ky > 0) v -= (Real)ky * ekp[1 * W::EKS + ((ky - 1) * (LD + 1) + jy) * (LCD + 2) + uu];
                dy[uu] = v;
              }
            });
            sfor<0, sz + 2>([&](auto vI) {
              constexpr int vv = decltype(vI)::value;
              ez[vv] = ekp[2 * W::EKS + (kz * (LD + 1) + jz) * (LCD + 2) + vv];
              {
                Real v = ga2 * ekp[2 * W::EKS + ((kz + 1) * (LD + 1) + jz) * (LCD + 2) + vv];
                if constexpr (kz > 0) v -= (Real)kz * ekp[2 * W::EKS + ((kz - 1) * (LD + 1) + jz) * (LCD + 2) + vv];
                dz[vv] = v;
              }
            });
            // h0 over t+u+v <= la+lb+1, hC (C-derivative) over t+u+v <= la+lb
            Real h0[NH1], hC[3][NH0];
            sfor<0, NH1>([&](auto kI) {
              constexpr int k = decltype(kI)::value;
              constexpr TUV e = herm_decode(k);
              constexpr bool low = k < NH0;
              Real s0 = (Real)0, cx = (Real)0, cy = (Real)0, cz = (Real)0;
              sfor<0, sx + 2>([&](auto tI) {
                constexpr int tt = decltype(tI)::value;
                sfor<0, sy + 2>([&](auto uI) {
                  constexpr int uu = decltype(uI)::value;
                  sfor<0, sz + 2>([&](auto vI) {
                    constexpr int vv = decltype(vI)::value;
                    constexpr int over = (tt > sx) + (uu > sy) + (vv > sz);
                    if constexpr (over <= 1 && (over == 0 || low)) {
                      Real r = R[herm_index(e.t + tt, e.u + uu, e.v + vv)];
                      if constexpr ((tt + uu + vv) & 1) r = -r;
                      if constexpr (over == 0) s0 += ex[tt] * ey[uu] * ez[vv] * r;
                      if constexpr (low) {
                        if constexpr (uu <= sy && vv <= sz) cx += dx[tt] * ey[uu] * ez[vv] * r;
                        if constexpr (tt <= sx && vv <= sz) cy += ex[tt] * dy[uu] * ez[vv] * r;
                        if constexpr (tt <= sx && uu <= sy) cz += ex[tt] * ey[uu] * dz[vv] * r;
                      }
                    }
                  });
                });
              });
              h0[k] = s0;
              if constexpr (low) {
                hC[0][k] = cx;
                hC[1][k] = cy;
                hC[2][k] = cz;
              }
            });
            // each bra pair ab against its A-derivative, P-shifted and plain E
            sfor<0, NA>([&](auto aI) {
              constexpr int a = decltype(aI)::value;
              constexpr int ax = cart_lx(LA, a), ay = cart_ly(LA, a), az = cart_lz(LA, a);
              sfor<0, NB>([&](auto bI) {
                constexpr int b = decltype(bI)::value;
                constexpr int bx = cart_lx(LB, b), by = cart_ly(LB, b), bz = cart_lz(LB, b);
                constexpr int rx = ax + bx, ry = ay + by, rz = az + bz;
                Real fx[rx + 2], fy[ry + 2], fz[rz + 2], gx[rx + 2], gy[ry + 2], gz[rz + 2];
                sfor<0, rx + 2>([&](auto tI) {
                  constexpr int t = decltype(tI)::value;
                  fx[t] = ebp[0 * W::EBS + (ax * (LB + 1) + bx) * (LAB + 2) + t];
                  {
                    Real v = al2 * ebp[0 * W::EBS + ((ax + 1) * (LB + 1) + bx) * (LAB + 2) + t];
                    if constexpr (ax > 0) v -= (Real)ax * ebp[0 * W::EBS + ((ax - 1) * (LB + 1) + bx) * (LAB + 2) + t];
                    gx[t] = v;
                  }
                });
                sfor<0, ry + 2>([&](auto uI) {
                  constexpr int u = decltype(uI)::value;
                  fy[u] = ebp[1 * W::EBS + (ay * (LB + 1) + by) * (LAB + 2) + u];
                  {
                    Real v = al2 * ebp[1 * W::EBS + ((ay + 1) * (LB + 1) + by) * (LAB + 2) + u];
                    if constexpr (ay > 0) v -= (Real)ay * ebp[1 * W::EBS + ((ay - 1) * (LB + 1) + by) * (LAB + 2) + u];
                    gy[u] = v;
                  }
                });
                sfor<0, rz + 2>([&](auto vI) {
                  constexpr int v = decltype(vI)::value;
                  fz[v] = ebp[2 * W::EBS + (az * (LB + 1) + bz) * (LAB + 2) + v];
                  {
                    Real w = al2 * ebp[2 * W::EBS + ((az + 1) * (LB + 1) + bz) * (LAB + 2) + v];
                    if constexpr (az > 0) w -= (Real)az * ebp[2 * W::EBS + ((az - 1) * (LB + 1) + bz) * (LAB + 2) + v];
                    gz[v] = w;
                  }
                });
                Real sm[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) sm[i] = (Real)0;
                sfor<0, rx + 2>([&](auto tI) {
                  constexpr int t = decltype(tI)::value;
                  sfor<0, ry + 2>([&](auto uI) {
                    constexpr int u = decltype(uI)::value;
                    sfor<0, rz + 2>([&](auto vI) {
                      constexpr int v = decltype(vI)::value;
                      constexpr int over = (t > rx) + (u > ry) + (v > rz);
                      if constexpr (over <= 1) {
                        const Real hv = h0[herm_index(t, u, v)];
                        if constexpr (u <= ry && v <= rz) sm[0] += gx[t] * fy[u] * fz[v] * hv;
                        if constexpr (t <= rx && v <= rz) sm[1] += fx[t] * gy[u] * fz[v] * hv;
                        if constexpr (t <= rx && u <= ry) sm[2] += fx[t] * fy[u] * gz[v] * hv;
                        if constexpr (over == 0) {
                          const Real e0 = fx[t] * fy[u] * fz[v];
                          // P = A + B: the bra distribution moved along x, y, z
                          sm[3] += e0 * h0[herm_index(t + 1, u, v)];
                          sm[4] += e0 * h0[herm_index(t, u + 1, v)];
                          sm[5] += e0 * h0[herm_index(t, u, v + 1)];
                          sm[6] += e0 * hC[0][herm_index(t, u, v)];
                          sm[7] += e0 * hC[1][herm_index(t, u, v)];
                          sm[8] += e0 * hC[2][herm_index(t, u, v)];
                        }
                      }
                    });
                  });
                });
                const Real gm = sh.gam[(a * NB + b) * NCD + c * ND + d] * pref;
#pragma unroll
                for (int i = 0; i < 9; ++i) acc[i] += gm * sm[i];
              });
            });
          }
        });
      }
    }
  }
  double accd[9];  // the wave's sums in FP64 whatever the arithmetic precision
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    accd[i] = (double)acc[i];
    for (int w = 32; w > 0; w >>= 1) accd[i] += __shfl_xor(accd[i], w);
  }
  if (lane == 0) {
    const double f = (args.sp_diag[bp] ? 1.0 : 2.0) * (args.sp_diag[kp] ? 1.0 : 2.0) *
                     (bp == kp ? 1.0 : 2.0) * 0.5;
    const int A = args.sp_atom[2 * bp], B = args.sp_atom[2 * bp + 1];
    const int C = args.sp_atom[2 * kp], D = args.sp_atom[2 * kp + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double gA = f * accd[k], gP = f * accd[3 + k], gC = f * accd[6 + k];
      atomicAdd(&args.grad[3 * A + k], gA);
      atomicAdd(&args.grad[3 * B + k], gP - gA);
      atomicAdd(&args.grad[3 * C + k], gC);
      atomicAdd(&args.grad[3 * D + k], -(gP + gC));
    }
  }
}

// ---------------------------------------------------------------------------
// The same one-lane-per-primitive-quartet layout with the bra side contracted
// with G first, as in the one-wave kernel (grad_impl.h): per bra primitive ib
// and ket pair cd of the launch's range, the wave builds in LDS
//   Y0[ib][cd][k]   = sum_ab G_abcd E^{ab}_k               (k < nherm(la+lb))
//   YA[ib][x][cd][k] = sum_ab G_abcd dE^{ab}_k / dA_x       (k < nherm(la+lb+1))
// and each lane's primitive quartet then needs only its h lists against Y:
//   dA_x += pref sum_k YA_x[k] h0[k]
//   dP_x += pref sum_k Y0[k] h0[k + 1_x]      (P = A + B, shift of R)
//   dC_x += pref sum_k Y0[k] hC_x[k]
// instead of every bra pair ab against its own E products (the work per
// primitive quartet no longer grows with the number of bra Cartesian pairs;
// Y is built once per bra primitive and shared by its ket primitives).
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename Real, int NCL>
struct GradYPrim {
  static constexpr int NH1 = nherm(LA + LB + 1), NH0 = nherm(LA + LB);
  struct YShared {
    Real y0[GMAXP][NCL][NH0];
    Real ya[GMAXP][3][NCL][NH1];
  };
};

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CD1>
__global__ __launch_bounds__(64) void grad2e_yprim_kernel(Grad2eArgs args) {
  using W = GradPrim<LA, LB, LC, LD, Real>;
  constexpr int LAB = W::LAB, LCD = W::LCD, L1 = W::L1;
  constexpr int NA = W::NA, NB = W::NB, NC = W::NC, ND = W::ND;
  constexpr int NAB = W::NAB, NCD = W::NCD, NH1 = W::NH1, NH0 = W::NH0;
  constexpr int NCL = CD1 - CD0;
  using Y = GradYPrim<LA, LB, LC, LD, Real, NCL>;
  __shared__ typename W::Shared sh;
  __shared__ typename Y::YShared ysh;

  const int lane = threadIdx.x;
  const int64_t gid = flat_block();
  if (gid >= args.nquartets) return;  // whole work-group (one wave)
  int lo = 0, hi = args.nbra - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.offsets[mid] <= gid)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int bp = args.bra_list[lo];
  const int kp = args.ket_list[args.kstart[lo] + (int)(gid - args.offsets[lo])];
  const Real* __restrict__ pp = gpairs<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];
  const int64_t oa = args.sp_aoc[2 * bp], ob = args.sp_aoc[2 * bp + 1];
  const int64_t oc = args.sp_aoc[2 * kp], od = args.sp_aoc[2 * kp + 1];
  const int64_t nc = args.ncart;

  // G block (the whole block: the density screen sees every cd)
  double gmax = 0.0;
  for (int e = lane; e < NAB * NCD; e += 64) {
    const int ab = e / NCD, cd = e % NCD;
    const int64_t mu = oa + ab / NB, nu = ob + ab % NB, la = oc + cd / ND, si = od + cd % ND;
    double k = args.dk0[mu * nc + la] * args.dk0[nu * nc + si] +
               args.dk0[mu * nc + si] * args.dk0[nu * nc + la];
    if (args.nk > 1)
      k += args.dk1[mu * nc + la] * args.dk1[nu * nc + si] +
           args.dk1[mu * nc + si] * args.dk1[nu * nc + la];
    const Real gv = (Real)(args.dj[mu * nc + nu] * args.dj[la * nc + si] - args.cx * k);
    sh.gam[e] = gv;
    gmax = fmax(gmax, fabs((double)gv));
  }
  if (args.tau_density > 0.0) {
    for (int w = 32; w > 0; w >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, w));
    if (gmax * args.schwarz[bp] * args.schwarz[kp] < args.tau_density) return;  // uniform
  }

  Real acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = (Real)0;

  for (int b0 = 0; b0 < npb; b0 += GMAXP) {
    const int nb = min(GMAXP, npb - b0);
    __syncthreads();  // the previous block's tables are no longer read
    if (lane < 3 * nb) {
      const int ib = lane / 3, dim = lane - 3 * ib;
      const Real* P = pp + (int64_t)(pb0 + b0 + ib) * PP_STRIDE;
      Real E[LA + 2][LB + 1][LAB + 2];
      hermite_e<LA + 1, LB, Real>(E, P[PP_PAX + dim], P[PP_PBX + dim], (Real)0.5 / P[PP_P]);
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
          for (int t = 0; t <= LAB + 1; ++t)
            sh.eb[ib][dim][(i * (LB + 1) + j) * (LAB + 2) + t] = E[i][j][t];
    }
    __syncthreads();
    // Y over (ib, cl, k): the item's E columns at (t, u, v) in registers, every
    // ab unrolled
    for (int it = lane; it < nb * NCL * NH1; it += 64) {
      const int ib = it / (NCL * NH1), r = it - ib * (NCL * NH1);
      const int cl = r / NH1, k = r - cl * NH1;
      const int tuv = kHermDecode<LAB + 1>.tuv[k];
      const int t = tuv & 255, u = (tuv >> 8) & 255, v = tuv >> 16;
      const Real al2 = (Real)2 * pp[(int64_t)(pb0 + b0 + ib) * PP_STRIDE + PP_ALPHA];
      Real ex[LA + 2][LB + 1], ey[LA + 2][LB + 1], ez[LA + 2][LB + 1];
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
          ex[i][j] = sh.eb[ib][0][(i * (LB + 1) + j) * (LAB + 2) + t];
          ey[i][j] = sh.eb[ib][1][(i * (LB + 1) + j) * (LAB + 2) + u];
          ez[i][j] = sh.eb[ib][2][(i * (LB + 1) + j) * (LAB + 2) + v];
        }
      Real y0 = (Real)0, yx = (Real)0, yy = (Real)0, yz = (Real)0;
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) {
        const int a = ab / NB, b = ab - a * NB;
        const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
        const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
        const Real g = sh.gam[ab * NCD + CD0 + cl];
        const Real e0x = ex[ax][bx], e0y = ey[ay][by], e0z = ez[az][bz];
        Real eax = al2 * ex[ax + 1][bx], eay = al2 * ey[ay + 1][by], eaz = al2 * ez[az + 1][bz];
        if (ax > 0) eax -= (Real)ax * ex[ax > 0 ? ax - 1 : 0][bx];
        if (ay > 0) eay -= (Real)ay * ey[ay > 0 ? ay - 1 : 0][by];
        if (az > 0) eaz -= (Real)az * ez[az > 0 ? az - 1 : 0][bz];
        const Real gx = g * e0x, gyz = g * (e0y * e0z);
        y0 += gyz * e0x;
        yx += gyz * eax;
        yy += gx * e0z * eay;
        yz += gx * e0y * eaz;
      }
      if (k < NH0) ysh.y0[ib][cl][k] = y0;  // zero for t + u + v = la + lb + 1
      ysh.ya[ib][0][cl][k] = yx;
      ysh.ya[ib][1][cl][k] = yy;
      ysh.ya[ib][2][cl][k] = yz;
    }

    for (int k0 = 0; k0 < npk; k0 += GMAXP) {
      const int nk = min(GMAXP, npk - k0);
      __syncthreads();  // Y written; the previous ket block's tables no longer read
      if (lane >= 32 && lane < 32 + 3 * nk) {
        const int ik = (lane - 32) / 3, dim = lane - 32 - 3 * ik;
        const Real* Q = pp + (int64_t)(pk0 + k0 + ik) * PP_STRIDE;
        Real E[LC + 2][LD + 1][LCD + 2];
        hermite_e<LC + 1, LD, Real>(E, Q[PP_PAX + dim], Q[PP_PBX + dim], (Real)0.5 / Q[PP_P]);
#pragma unroll
        for (int i = 0; i <= LC + 1; ++i)
#pragma unroll
          for (int j = 0; j <= LD; ++j)
#pragma unroll
            for (int t = 0; t <= LCD + 1; ++t)
              sh.ek[ik][dim][(i * (LD + 1) + j) * (LCD + 2) + t] = E[i][j][t];
      }
      __syncthreads();

      for (int q = lane; q < nb * nk; q += 64) {
        const int ib = q / nk, ik = q - ib * nk;
        const Real* P = pp + (int64_t)(pb0 + b0 + ib) * PP_STRIDE;
        const Real* Q = pp + (int64_t)(pk0 + k0 + ik) * PP_STRIDE;
        const Real p = P[PP_P], qq = Q[PP_P];
        const Real ga2 = (Real)2 * Q[PP_ALPHA];
        const Real pqm = p * qq, ps = p + qq, alpha = pqm / ps;
        const Real X = P[PP_PX] - Q[PP_PX], Yc = P[PP_PX + 1] - Q[PP_PX + 1],
                   Z = P[PP_PX + 2] - Q[PP_PX + 2];
        Real F[L1 + 1];
        boys_eval<L1, Real>(alpha * (X * X + Yc * Yc + Z * Z), args.boys64, F);
        Real R[nherm(L1)];
        hermite_r_tet<L1, Real>(R, F, (Real)-2 * alpha, X, Yc, Z);
        const Real pref = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * P[PP_K] * Q[PP_K];
        const Real* __restrict__ ekp = &sh.ek[ik][0][0];
        Real sm[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) sm[i] = (Real)0;
        sfor<CD0, CD1>([&](auto cdI) {
          constexpr int cd = decltype(cdI)::value, cl = cd - CD0;
          constexpr int c = cd / ND, d = cd % ND;
          constexpr int kx = cart_lx(LC, c), ky = cart_ly(LC, c), kz = cart_lz(LC, c);
          constexpr int jx = cart_lx(LD, d), jy = cart_ly(LD, d), jz = cart_lz(LD, d);
          constexpr int sx = kx + jx, sy = ky + jy, sz = kz + jz;
          // this ket pair's E (and C-derivative E) factors per direction
          Real ex[sx + 2], ey[sy + 2], ez[sz + 2], dx[sx + 2], dy[sy + 2], dz[sz + 2];
          sfor<0, sx + 2>([&](auto tI) {
            constexpr int tt = decltype(tI)::value;
            ex[tt] = ekp[0 * W::EKS + (kx * (LD + 1) + jx) * (LCD + 2) + tt];
            Real v = ga2 * ekp[0 * W::EKS + ((kx + 1) * (LD + 1) + jx) * (LCD + 2) + tt];
            if constexpr (kx > 0) v -= (Real)kx * ekp[0 * W::EKS + ((kx - 1) * (LD + 1) + jx) * (LCD + 2) + tt];
            dx[tt] = v;
          });
          sfor<0, sy + 2>([&](auto uI) {
            constexpr int uu = decltype(uI)::value;
            ey[uu] = ekp[1 * W::EKS + (ky * (LD + 1) + jy) * (LCD + 2) + uu];
            Real v = ga2 * ekp[1 * W::EKS + ((ky + 1) * (LD + 1) + jy) * (LCD + 2) + uu];
            if constexpr (ky > 0) v -= (Real)ky * ekp[1 * W::EKS + ((ky - 1) * (LD + 1) + jy) * (LCD + 2) + uu];
            dy[uu] = v;
          });
          sfor<0, sz + 2>([&](auto vI) {
            constexpr int vv = decltype(vI)::value;
            ez[vv] = ekp[2 * W::EKS + (kz * (LD + 1) + jz) * (LCD + 2) + vv];
            Real v = ga2 * ekp[2 * W::EKS + ((kz + 1) * (LD + 1) + jz) * (LCD + 2) + vv];
            if constexpr (kz > 0) v -= (Real)kz * ekp[2 * W::EKS + ((kz - 1) * (LD + 1) + jz) * (LCD + 2) + vv];
            dz[vv] = v;
          });
          // h0[k] (t+u+v <= la+lb+1) and hC_x,y,z[k] (t+u+v <= la+lb), each
          // taken against Y as soon as it is formed (no h list is kept):
          // dA_x: YA_x[k] h0[k]; dP_x: Y0[k - 1_x] h0[k] (the shift of R read
          // from the other side); dC_x: Y0[k] hC_x[k]
          const Real* __restrict__ ya0 = &ysh.ya[ib][0][cl][0];
          const Real* __restrict__ ya1 = &ysh.ya[ib][1][cl][0];
          const Real* __restrict__ ya2 = &ysh.ya[ib][2][cl][0];
          const Real* __restrict__ y0p = &ysh.y0[ib][cl][0];
          sfor<0, NH1>([&](auto kI) {
            constexpr int k = decltype(kI)::value;
            constexpr TUV e = herm_decode(k);
            constexpr bool low = k < NH0;
            Real s0 = (Real)0, cx = (Real)0, cy = (Real)0, cz = (Real)0;
            sfor<0, sx + 2>([&](auto tI) {
              constexpr int tt = decltype(tI)::value;
              sfor<0, sy + 2>([&](auto uI) {
                constexpr int uu = decltype(uI)::value;
                sfor<0, sz + 2>([&](auto vI) {
                  constexpr int vv = decltype(vI)::value;
                  constexpr int over = (tt > sx) + (uu > sy) + (vv > sz);
                  if constexpr (over <= 1 && (over == 0 || low)) {
                    Real r = R[herm_index(e.t + tt, e.u + uu, e.v + vv)];
                    if constexpr ((tt + uu + vv) & 1) r = -r;
                    if constexpr (over == 0) s0 += ex[tt] * ey[uu] * ez[vv] * r;
                    if constexpr (low) {
                      if constexpr (uu <= sy && vv <= sz) cx += dx[tt] * ey[uu] * ez[vv] * r;
                      if constexpr (tt <= sx && vv <= sz) cy += ex[tt] * dy[uu] * ez[vv] * r;
                      if constexpr (tt <= sx && uu <= sy) cz += ex[tt] * ey[uu] * dz[vv] * r;
                    }
                  }
                });
              });
            });
            sm[0] += ya0[k] * s0;
            sm[1] += ya1[k] * s0;
            sm[2] += ya2[k] * s0;
            if constexpr (e.t > 0) sm[3] += y0p[herm_index(e.t - 1, e.u, e.v)] * s0;
            if constexpr (e.u > 0) sm[4] += y0p[herm_index(e.t, e.u - 1, e.v)] * s0;
            if constexpr (e.v > 0) sm[5] += y0p[herm_index(e.t, e.u, e.v - 1)] * s0;
            if constexpr (low) {
              const Real y = y0p[k];
              sm[6] += y * cx;
              sm[7] += y * cy;
              sm[8] += y * cz;
            }
          });
        });
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += pref * sm[i];
      }
    }
  }
  double accd[9];  // the wave's sums in FP64 whatever the arithmetic precision
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    accd[i] = (double)acc[i];
    for (int w = 32; w > 0; w >>= 1) accd[i] += __shfl_xor(accd[i], w);
  }
  if (lane == 0) {
    const double f = (args.sp_diag[bp] ? 1.0 : 2.0) * (args.sp_diag[kp] ? 1.0 : 2.0) *
                     (bp == kp ? 1.0 : 2.0) * 0.5;
    const int A = args.sp_atom[2 * bp], B = args.sp_atom[2 * bp + 1];
    const int C = args.sp_atom[2 * kp], D = args.sp_atom[2 * kp + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double gA = f * accd[k], gP = f * accd[3 + k], gC = f * accd[6 + k];
      atomicAdd(&args.grad[3 * A + k], gA);
      atomicAdd(&args.grad[3 * B + k], gP - gA);
      atomicAdd(&args.grad[3 * C + k], gC);
      atomicAdd(&args.grad[3 * D + k], -(gP + gC));
    }
  }
}

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int NCD, int CH>
void grad2e_launch_yprim_chunks(dim3 grid, const Grad2eArgs& a, hipStream_t s) {
  if constexpr (CD0 < NCD) {
    constexpr int CD1 = CD0 + CH < NCD ? CD0 + CH : NCD;
    hipLaunchKernelGGL((grad2e_yprim_kernel<LA, LB, LC, LD, Real, CD0, CD1>), grid, dim3(64), 0, s, a);
    grad2e_launch_yprim_chunks<LA, LB, LC, LD, Real, CD1, NCD, CH>(grid, a, s);
  }
}

template <int LA, int LB, int LC, int LD, typename Real = double, int CH = 1>
hipError_t grad2e_launch_yprim(const Grad2eArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  grad2e_launch_yprim_chunks<LA, LB, LC, LD, Real, 0, ncart(LC) * ncart(LD), CH>(
      grid_1d(a.nquartets, 64), a, s);
  return hipGetLastError();
}

// The launches of ket pairs [CDA, CDB) only (CH per launch): the build
// spreads the heavy classes' instantiations over several translation units.
template <int LA, int LB, int LC, int LD, typename Real, int CH, int CDA, int CDB>
hipError_t grad2e_launch_yprim_range(const Grad2eArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  grad2e_launch_yprim_chunks<LA, LB, LC, LD, Real, CDA, CDB, CH>(grid_1d(a.nquartets, 64), a, s);
  return hipGetLastError();
}

// CH ket Cartesian pairs cd per launch, each launch adding its share of the
// gradient (a lane holds R and the h lists of one cd at a time either way).
// Every launch repeats the quartet's G block, E tables, Boys values and R,
// so fewer launches pay where the larger unrolled body still fits the
// registers; the build picks CH per class and precision by measurement.
template <int LA, int LB, int LC, int LD, typename Real, int CD0, int NCD, int CH>
void grad2e_launch_prim_chunks(dim3 grid, const Grad2eArgs& a, hipStream_t s) {
  if constexpr (CD0 < NCD) {
    constexpr int CD1 = CD0 + CH < NCD ? CD0 + CH : NCD;
    hipLaunchKernelGGL((grad2e_prim_kernel<LA, LB, LC, LD, Real, CD0, CD1>), grid, dim3(64), 0, s, a);
    grad2e_launch_prim_chunks<LA, LB, LC, LD, Real, CD1, NCD, CH>(grid, a, s);
  }
}

template <int LA, int LB, int LC, int LD, typename Real = double, int CH = 1>
hipError_t grad2e_launch_prim(const Grad2eArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  grad2e_launch_prim_chunks<LA, LB, LC, LD, Real, 0, ncart(LC) * ncart(LD), CH>(
      grid_1d(a.nquartets, 64), a, s);
  return hipGetLastError();
}

}  // namespace qcmd
