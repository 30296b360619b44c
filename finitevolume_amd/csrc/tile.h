// Cell tiles in LDS for the fused Euler stage: a tile is a range of
// graph-adjacent cells (the owned cells) with two halo rings; its states and
// primitives, and the MUSCL increments of its faces, are staged in LDS
// structure-of-arrays, so every neighbour access inside a tile is an LDS read.
#pragma once

#include "physics.h"
#include "solver.h"

namespace fvk {

// LDS views of one tile (structure of arrays: consecutive local cells or
// faces are consecutive 8-byte words, so a wavefront's own-cell accesses
// do not conflict on the banks)
struct TileLds {
    double* U;   // [4][MG] conserved state, owned cells then ring 1
    double* W;   // [3][MG] primitives u, v, p of the owned and ring-1 cells (second order; rho is U's row 0)
    double* W2;  // [4][M2] primitives rho, u, v, p of ring 2 (second order)
    double* I;   // [8][MF] second order: per tile face, the primitive increments G . r of cell_left
                 //   (rows 0-3) and cell_right (rows 4-7), written by the gradients of the two cells
    double* F;   // [4][MF] face flux x length (second order: over I's rows 0-3, each face's own column)
    double* Dt;  // [MF] dt candidates (second order: over I's row 4)
    int MG, M2, MF;
    int ng;      // this tile's owned + ring-1 cells: local indices below it are in W, the rest in W2
};

// limited LSQ gradient of local cell j (as the per-cell gradient kernel),
// from LDS; then, for each slot whose face is a tile face, the MUSCL
// increments G . r of the face's reconstruction on this cell's side (as
// reconstruct(): r the offset to the face midpoint, frec's value; rec,
// stride MG, holds it on a periodic mesh, else egeo's equal midpoint
// offset).  f0: the tile's first face in the tile face arrays.
//
// EN(k) and EG(k) give slot k's local ELL entry and tile geometry; EG is
// called again in each of the three slot loops.  With arrays held in
// registers (tile_grad) they are register reads; tile_grad_lean reads them
// from memory in each loop instead, so wide cells (Voronoi polygons, up to
// nine slots) do not hold nine double4 across the function.
template <int D, bool LEAN, class EnF, class EgF>
__device__ __forceinline__ void tile_grad_with(const MeshView& m, double g, int limiter, const TileLds& L, int j,
                                               EnF EN, EgF EG, const double4& iv, int f0, const double2* rec) {
    int en[D];
#pragma unroll
    for (int k = 0; k < D; ++k) en[k] = EN(k);
    const int MG = L.MG;
    const State Uc{L.U[j], L.U[MG + j], L.U[2 * MG + j], L.U[3 * MG + j]};
    const double wc[4] = {L.U[j], L.W[j], L.W[MG + j], L.W[2 * MG + j]};
    double bx[4] = {0.0, 0.0, 0.0, 0.0}, by[4] = {0.0, 0.0, 0.0, 0.0};
    double dlo[4] = {0.0, 0.0, 0.0, 0.0}, dhi[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (en[k] == -1) continue;
        const double4 egk = EG(k);
        double wn[4];
        if (tell_is_boundary(en[k])) {  // the ghost state; the face from the tile's list or, ring 1, the mesh
            const int x = tell_index(en[k]);
            const bool ring1 = en[k] & 1;
            const double4 fg = ring1 ? m.fgeo[x] : m.tfgeo[f0 + x];
            const DevBC& bc = m.bc[ring1 ? -1 - m.fr[x] : -1 - tface_right(m.tfc[f0 + x])];
            const State Gs = euler_ghost_state(bc.type, bc.ff, Uc, fg.x, fg.y, g);
            const double wg0 = Gs.r, wg1 = fdiv(Gs.ru, Gs.r), wg2 = fdiv(Gs.rv, Gs.r), wg3 = pressure(Gs, g);
            wn[0] = 0.5 * (wc[0] + wg0);
            wn[1] = 0.5 * (wc[1] + wg1);
            wn[2] = 0.5 * (wc[2] + wg2);
            wn[3] = 0.5 * (wc[3] + wg3);
        } else {
            const int l = tell_nb(en[k]);
            if (l < L.ng) {
                wn[0] = L.U[l];
                wn[1] = L.W[l];
                wn[2] = L.W[MG + l];
                wn[3] = L.W[2 * MG + l];
            } else {
                const double* w = L.W2 + (l - L.ng);
                const int sw = L.M2;
                wn[0] = w[0];
                wn[1] = w[sw];
                wn[2] = w[2 * sw];
                wn[3] = w[3 * sw];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double dphi = wn[q] - wc[q];
            bx[q] += egk.x * dphi;
            by[q] += egk.y * dphi;
            venkat_extend(dlo[q], dhi[q], dphi);
        }
    }
    double gx[4], gy[4], lim[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gx[q] = iv.x * bx[q] + iv.y * by[q];
        gy[q] = iv.y * bx[q] + iv.z * by[q];
        lim[q] = 1.0;
    }
    if (limiter == FVK_LIMITER_VENKATAKRISHNAN) {
        const double eps2 = iv.w;
        VenkatMin vm[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) vm[q] = VenkatMin{1.0, 1.0};
        if (LEAN) __asm__ volatile("" ::: "memory");  // reload the slot geometry below
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k] == -1) continue;
            const double4 egk = EG(k);
            const double rx = egk.z, ry = egk.w;
#pragma unroll
            for (int q = 0; q < 4; ++q) venkat_pick_d(vm[q], dlo[q], dhi[q], gx[q] * rx + gy[q] * ry, eps2);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) lim[q] = venkat_value(vm[q]);
    } else if (limiter == FVK_LIMITER_FIRST_ORDER) {
#pragma unroll
        for (int q = 0; q < 4; ++q) lim[q] = 0.0;
    }
    double G[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        G[2 * q] = lim[q] * gx[q];
        G[2 * q + 1] = lim[q] * gy[q];
    }
    const int MF = L.MF;
    if (LEAN) __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int lf = tell_lf(en[k]);
        if (en[k] == -1 || lf < 0) continue;
        double rx, ry;
        if (rec) {
            const double2 r = rec[k * MG];
            rx = r.x;
            ry = r.y;
        } else {
            const double4 egk = EG(k);
            rx = egk.z;
            ry = egk.w;
        }
        double* I = L.I + (tell_owner(en[k]) ? 0 : 4 * MF) + lf;
        I[0] = G[0] * rx + G[1] * ry;
        I[MF] = G[2] * rx + G[3] * ry;
        I[2 * MF] = G[4] * rx + G[5] * ry;
        I[3 * MF] = G[6] * rx + G[7] * ry;
    }
}

template <int D>
__device__ __forceinline__ void tile_grad(const MeshView& m, double g, int limiter, const TileLds& L, int j,
                                          const int (&en)[D], const double4 (&eg)[D], const double4& iv, int f0,
                                          const double2* rec) {
    tile_grad_with<D, false>(
        m, g, limiter, L, j, [&](int k) { return en[k]; }, [&](int k) { return eg[k]; }, iv, f0, rec);
}

// te, tg: the cell's first slot in the tile ELL and tile geometry (stride MG)
template <int D>
__device__ __forceinline__ void tile_grad_lean(const MeshView& m, double g, int limiter, const TileLds& L, int j,
                                               const int* te, const double4* tg, const double4& iv, int f0,
                                               const double2* rec) {
    const int MG = L.MG;
    tile_grad_with<D, true>(
        m, g, limiter, L, j, [&](int k) { return te[k * MG]; }, [&](int k) { return tg[k * MG]; }, iv, f0, rec);
}

// stage local cell j: conserved state (owned and ring-1 cells) and, for
// second order, its primitives (W for owned and ring 1, W2 for ring 2)
__device__ __forceinline__ void tile_stage(const TileLds& L, int j, const double4& u, double g, bool second_order) {
    const int MG = L.MG, ng = L.ng;
    if (j < ng) {
        L.U[j] = u.x;
        L.U[MG + j] = u.y;
        L.U[2 * MG + j] = u.z;
        L.U[3 * MG + j] = u.w;
    }
    if (second_order) {
        const State U{u.x, u.y, u.z, u.w};
        const double pu = fdiv(U.ru, U.r), pv = fdiv(U.rv, U.r), pp = pressure(U, g);
        if (j < ng) {
            L.W[j] = pu;
            L.W[MG + j] = pv;
            L.W[2 * MG + j] = pp;
        } else {
            double* w = L.W2 + (j - ng);
            const int sw = L.M2;
            w[0] = U.r;
            w[sw] = pu;
            w[2 * sw] = pv;
            w[3 * sw] = pp;
        }
    }
}

}  // namespace fvk
