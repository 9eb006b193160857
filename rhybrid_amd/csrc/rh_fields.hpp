// rh_fields.hpp -- device kernels of propagateB (src/hybrid_propagator.cpp:
// 42-440), variant USE_NODE_UE + USE_EDGE_J (src/Makefile:1-7).
//
// Grid arrays are structure of arrays: a vector field is one allocation of
// 3*nc with component l at [l*nc + n].  One thread per cell; neighbours are
// addressed through the periodic/missing topology of rh_geom.hpp, so every
// kernel keeps the reference's rule for which cells it updates:
//   cell-centred kernels only in cells with all 26 neighbours;
//   node/face kernels skip cells missing a +x/+y/+z neighbour.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_geom.hpp"
#include "rh_particles.hpp"

namespace rh {

template <typename Real>
struct V3 {
    Real *x, *y, *z;
    __device__ __forceinline__ Real& c(int l, int n) const { return l == 0 ? x[n] : (l == 1 ? y[n] : z[n]); }
};

// 3D launch: blocks of 32x8 cells in x,y, one z layer per grid row
#define RH_CELL_IJK                                      \
    const int i = blockIdx.x * 32 + threadIdx.x;         \
    const int j = blockIdx.y * 8 + threadIdx.y;          \
    const int k = blockIdx.z;                            \
    const bool in_grid = i < g.nx && j < g.ny;           \
    const int n = in_grid ? cell_index(g, i, j, k) : 0;

#define RH_CELL_PROLOGUE \
    RH_CELL_IJK          \
    if (!in_grid) return;

// face2Cell (:1043-1056) at cell (i, j, k) = n
template <typename Real>
__device__ __forceinline__ void face2cell_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& face,
                                             Real& bx, Real& by, Real& bz) {
    bx = Real(0.5) * (face.x[n] + face.x[nbr(g, i, j, k, -1, 0, 0)]);
    by = Real(0.5) * (face.y[n] + face.y[nbr(g, i, j, k, 0, -1, 0)]);
    bz = Real(0.5) * (face.z[n] + face.z[nbr(g, i, j, k, 0, 0, -1)]);
}

// cell2Node (:1059-1089): node = mean of the 8 cells around the upper corner
template <typename Real>
__global__ __launch_bounds__(256) void k_cell2node(Geom g, const Real* __restrict__ cell, Real* __restrict__ node, int dim) {
    RH_CELL_PROLOGUE
    if (upper_missing(g, i, j, k)) return;
    const int c001 = nbr(g, i, j, k, 0, 0, 1), c010 = nbr(g, i, j, k, 0, 1, 0), c100 = nbr(g, i, j, k, 1, 0, 0);
    const int c011 = nbr(g, i, j, k, 0, 1, 1), c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1);
    const int c111 = nbr(g, i, j, k, 1, 1, 1);
    for (int l = 0; l < dim; ++l) {
        const Real* a = cell + (size_t)l * g.nc;
        node[(size_t)l * g.nc + n] =
            Real(0.125) * (a[n] + a[c001] + a[c010] + a[c100] + a[c011] + a[c110] + a[c101] + a[c111]);
    }
}

// cell2Node of several arrays in one launch (k_node_J)
template <typename Real>
struct Cell2NodeSet {
    const Real* from[3];
    Real* to[3];
    int dim[3];
    int n;
};
// neumannCell (:443-655).  Walls, then edges, then corners overwrite one
// another, so an exterior cell ends with the value of the neighbour lying
// opposite to all of its missing directions; that neighbour is never an
// exterior cell, which makes the cells independent.
__device__ __forceinline__ int neumann_source(const Geom& g, int i, int j, int k) {
    const int sx = has_neg(i, g.per[0]) ? (has_pos(i, g.nx, g.per[0]) ? 0 : -1) : 1;
    const int sy = has_neg(j, g.per[1]) ? (has_pos(j, g.ny, g.per[1]) ? 0 : -1) : 1;
    const int sz = has_neg(k, g.per[2]) ? (has_pos(k, g.nz, g.per[2]) ? 0 : -1) : 1;
    return nbr(g, i, j, k, sx, sy, sz);
}

template <typename Real>
__global__ __launch_bounds__(256) void k_neumann_cell(Geom g, const int* __restrict__ ext, int next,
                                                     Real* __restrict__ D, int dim) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const int src = neumann_source(g, i, j, k);
    if (src < 0) return;
    for (int l = 0; l < dim; ++l) D[(size_t)l * g.nc + n] = D[(size_t)l * g.nc + src];
}

// neumannFace (:659-908): per-component copy rules in the reference's order;
// for each component the last applicable rule wins.
struct FaceRule { uint8_t need; int8_t comp, a, b, c; };
// need bits: 1 -x, 2 +x, 4 -y, 8 +y, 16 -z, 32 +z (neighbours that must be missing)
constexpr FaceRule kFaceRules[48] = {
    {1, 1, 1, 0, 0}, {1, 2, 1, 0, 0}, {8, 0, 0, -1, 0}, {8, 2, 0, -1, 0},
    {4, 0, 0, 1, 0}, {4, 2, 0, 1, 0}, {32, 0, 0, 0, -1}, {32, 1, 0, 0, -1},
    {16, 0, 0, 0, 1}, {16, 1, 0, 0, 1}, {2, 1, -1, 0, 0}, {2, 2, -1, 0, 0},
    {1 | 4, 0, 0, 1, 0}, {1 | 4, 1, 1, 0, 0}, {1 | 4, 2, 1, 1, 0},
    {1 | 8, 0, 0, -1, 0}, {1 | 8, 2, 1, -1, 0},
    {2 | 4, 1, -1, 0, 0}, {2 | 4, 2, -1, 1, 0},
    {2 | 8, 2, -1, -1, 0},
    {1 | 16, 0, 0, 0, 1}, {1 | 16, 1, 1, 0, 1}, {1 | 16, 2, 1, 0, 0},
    {1 | 32, 0, 0, 0, -1}, {1 | 32, 1, 1, 0, -1},
    {2 | 16, 1, -1, 0, 1}, {2 | 16, 2, -1, 0, 0},
    {2 | 32, 1, -1, 0, -1},
    {4 | 16, 0, 0, 1, 1}, {4 | 16, 1, 0, 0, 1}, {4 | 16, 2, 0, 1, 0},
    {4 | 32, 0, 0, 1, -1}, {4 | 32, 1, 0, 0, -1},
    {8 | 16, 0, 0, -1, 1}, {8 | 16, 2, 0, -1, 0},
    {8 | 32, 0, 0, -1, -1},
    {1 | 4 | 16, 0, 0, 1, 1}, {1 | 4 | 16, 1, 1, 0, 1}, {1 | 4 | 16, 2, 1, 1, 0},
    {1 | 4 | 32, 0, 0, 1, -1}, {1 | 4 | 32, 1, 1, 0, -1},
    {1 | 8 | 16, 0, 0, -1, 1}, {1 | 8 | 16, 2, 1, -1, 0},
    {1 | 8 | 32, 0, 0, -1, -1},
    {2 | 4 | 16, 1, -1, 0, 1}, {2 | 4 | 16, 2, -1, 1, 0},
    {2 | 4 | 32, 1, -1, 0, -1},
    {2 | 8 | 16, 2, -1, -1, 0},
};

// The rule that wins for component l of a cell with missing-neighbour mask
// miss (the last applicable one), as the source offset (a, b, c) and a
// found flag: 64 masks x 3 components, built at compile time from the rules.
struct FaceSrc { int8_t a, b, c, found; };
struct FaceSrcTable { FaceSrc t[64][3]; };
constexpr FaceSrcTable make_face_src() {
    FaceSrcTable T{};
    for (int miss = 0; miss < 64; ++miss)
        for (int l = 0; l < 3; ++l) {
            T.t[miss][l] = FaceSrc{0, 0, 0, 0};
            for (int r = 47; r >= 0; --r) {
                const FaceRule R = kFaceRules[r];
                if (R.comp != l || (miss & R.need) != R.need) continue;
                T.t[miss][l] = FaceSrc{R.a, R.b, R.c, 1};
                break;
            }
        }
    return T;
}
__constant__ FaceSrcTable c_face_src = make_face_src();

__device__ __forceinline__ int missing_mask(const Geom& g, int i, int j, int k) {
    int m = 0;
    if (!has_neg(i, g.per[0])) m |= 1;
    if (!has_pos(i, g.nx, g.per[0])) m |= 2;
    if (!has_neg(j, g.per[1])) m |= 4;
    if (!has_pos(j, g.ny, g.per[1])) m |= 8;
    if (!has_neg(k, g.per[2])) m |= 16;
    if (!has_pos(k, g.nz, g.per[2])) m |= 32;
    return m;
}

template <typename Real>
__global__ __launch_bounds__(256) void k_neumann_face(Geom g, const int* __restrict__ ext, int next, V3<Real> F) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const int miss = missing_mask(g, i, j, k);
    for (int l = 0; l < 3; ++l) {
        const FaceSrc R = c_face_src.t[miss][l];
        if (!R.found) continue;
        const int src = nbr(g, i, j, k, R.a, R.b, R.c);
        if (src >= 0) F.c(l, n) = F.c(l, src);
    }
}

// setIMFCell / setIMFFace (:910-1040); wall order +x -x +y -y +z -z
template <typename Real>
__device__ __forceinline__ void imf_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& D, int face,
                                       int flags, Real bx, Real by, Real bz) {
    const int miss = missing_mask(g, i, j, k);
    // missing-mask bit of each wall in the +x -x +y -y +z -z order
    const int wall_bit[6] = {2, 1, 8, 4, 32, 16};
    for (int w = 0; w < 6; ++w) {
        if (!((flags >> w) & 1) || !(miss & wall_bit[w])) continue;
        const int axis = w / 2;
        if (!face || axis != 0) D.x[n] = bx;
        if (!face || axis != 1) D.y[n] = by;
        if (!face || axis != 2) D.z[n] = bz;
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_imf(Geom g, const int* __restrict__ ext, int next, V3<Real> D,
                                            int face, int flags, Real bx, Real by, Real bz) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= next) return;
    const int n = ext[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    imf_at(g, i, j, k, n, D, face, flags, bx, by, bz);
}

// face2Cell, then the exterior cells of cellB, cellJi and cellRhoQi by
// neumannCell, then setIMFCell on cellB (:122-155), in one launch.  An
// exterior cell's Neumann source is never exterior, so its face2Cell value is
// computed from faceB here as the launch's own thread for it computes it;
// cellJi and cellRhoQi are not written by this launch.
template <typename Real>
__global__ __launch_bounds__(256) void k_cell_B_bc(Geom g, V3<Real> face, V3<Real> cellB, V3<Real> Ji,
                                                  Real* __restrict__ rho, int imf_flags, Real bx, Real by, Real bz) {
    RH_CELL_PROLOGUE
    if (all_nbrs(g, i, j, k)) {
        Real cx, cy, cz;
        face2cell_at(g, i, j, k, n, face, cx, cy, cz);
        cellB.x[n] = cx; cellB.y[n] = cy; cellB.z[n] = cz;
        return;
    }
    const int src = neumann_source(g, i, j, k);
    if (src >= 0) {
        int si, sj, sk;
        cell_ijk(g, src, si, sj, sk);
        Real cx, cy, cz;
        face2cell_at(g, si, sj, sk, src, face, cx, cy, cz);
        cellB.x[n] = cx; cellB.y[n] = cy; cellB.z[n] = cz;
        Ji.x[n] = Ji.x[src]; Ji.y[n] = Ji.y[src]; Ji.z[n] = Ji.z[src];
        rho[n] = rho[src];
    }
    if (imf_flags) imf_at(g, i, j, k, n, cellB, 0, imf_flags, bx, by, bz);
}

// calcNodeJ (:1449-1522): nodeJ = curl(faceB)/mu0 on edges averaged to the
// node, divided by the whistler-limiter factor d.
template <typename Real>
__device__ __forceinline__ bool node_J_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& F, Real nbx,
                                          Real nby, Real nbz, Real rho_n, Real max_vw, Real (&J)[3]) {
    const int c100 = nbr(g, i, j, k, 1, 0, 0), c010 = nbr(g, i, j, k, 0, 1, 0), c001 = nbr(g, i, j, k, 0, 0, 1);
    const int c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1), c011 = nbr(g, i, j, k, 0, 1, 1);
    const Real ex1 = -F.z[n] + F.y[n] + F.z[c010] - F.y[c001];
    const Real ex2 = -F.z[c100] + F.y[c100] + F.z[c110] - F.y[c101];
    const Real ey1 = +F.z[n] + F.x[c001] - F.z[c100] - F.x[n];
    const Real ey2 = +F.z[c010] + F.x[c011] - F.z[c110] - F.x[c010];
    const Real ez1 = -F.y[n] + F.x[n] + F.y[c100] - F.x[c010];
    const Real ez2 = -F.y[c001] + F.x[c001] + F.y[c101] - F.x[c011];
    const Real e = Real(1.60217653e-19);
    const Real mu0 = Real(4.0e-7 * 3.14159265358979323846);
    Real d = Real(1), vw = Real(0);
    bool limited = false;
    const Real ne = rho_n / e;
    const Real Btot = sqrt(nbx * nbx + nby * nby + nbz * nbz);
    const Real dx = (Real)g.dx;
    if (ne > Real(0) && dx > Real(0))
        vw = Real(2) * Btot * Real(3.14159265358979323846) / (mu0 * ne * e * dx);
    if (vw > max_vw && max_vw > Real(0)) { d = vw / max_vw; limited = true; }
    const Real a = Real(0.5) / (dx * mu0 * d);
    J[0] = (ex1 + ex2) * a;
    J[1] = (ey1 + ey2) * a;
    J[2] = (ez1 + ez2) * a;
    return limited;
}

// cell2Node of cellB, cellRhoQi and cellJi, then calcNodeJ (:169-225) in one
// launch: both run on the same nodes, and calcNodeJ reads the node's own
// nodeB and nodeRhoQi (and faceB, after its Neumann and IMF copies, which
// come first: they do not touch the cell arrays cell2Node reads).
template <typename Real>
__global__ __launch_bounds__(256) void k_node_J(Geom g, Cell2NodeSet<Real> set, V3<Real> F, V3<Real> nodeJ,
                                               Real max_vw, unsigned long long* counters) {
    RH_CELL_IJK
    bool limited = false;
    if (in_grid && !upper_missing(g, i, j, k)) {
        const int c001 = nbr(g, i, j, k, 0, 0, 1), c010 = nbr(g, i, j, k, 0, 1, 0), c100 = nbr(g, i, j, k, 1, 0, 0);
        const int c011 = nbr(g, i, j, k, 0, 1, 1), c110 = nbr(g, i, j, k, 1, 1, 0), c101 = nbr(g, i, j, k, 1, 0, 1);
        const int c111 = nbr(g, i, j, k, 1, 1, 1);
        Real v[7];
        int q = 0;
        for (int s = 0; s < set.n; ++s)
            for (int l = 0; l < set.dim[s]; ++l, ++q) {
                const Real* a = set.from[s] + (size_t)l * g.nc;
                v[q] = Real(0.125) * (a[n] + a[c001] + a[c010] + a[c100] + a[c011] + a[c110] + a[c101] + a[c111]);
                set.to[s][(size_t)l * g.nc + n] = v[q];
            }
        // set = {cellB -> nodeB (3), cellRhoQi -> nodeRhoQi (1), cellJi -> nodeJi (3)}
        Real J[3];
        limited = node_J_at(g, i, j, k, n, F, v[0], v[1], v[2], v[3], max_vw, J);
        nodeJ.x[n] = J[0]; nodeJ.y[n] = J[1]; nodeJ.z[n] = J[2];
    }
    count_event(counters, C_MAX_VW, limited);
}

// calcCellUe (:1342-1387) and calcNodeUe (:1524-1584; NODE clamps rho in place)
template <typename Real, bool NODE>
__device__ __forceinline__ void ue_at(int n, const Real (&jj)[3], const V3<Real>& Ji, Real* __restrict__ rho,
                                      const uint8_t* __restrict__ inner, int use_hall, Real ox_, Real oy_, Real oz_,
                                      Real max_ue2, Real min_rho, Real (&u)[3], bool& c_ue, bool& c_rho) {
    if (inner[n]) {
        u[0] = ox_; u[1] = oy_; u[2] = oz_;
        return;
    }
    Real r = rho[n];
    if (NODE && r < min_rho) { r = min_rho; rho[n] = r; c_rho = true; }
    const Real ji[3] = {Ji.x[n], Ji.y[n], Ji.z[n]};
    for (int l = 0; l < 3; ++l) {
        if (fabs(r) > Real(0)) u[l] = use_hall ? -(jj[l] - ji[l]) / r : ji[l] / r;
        else u[l] = Real(0);
    }
    const Real u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (u2 > max_ue2) {
        const Real norm = sqrt(max_ue2 / u2);
        u[0] *= norm; u[1] *= norm; u[2] *= norm;
        c_ue = true;
    }
}

// node2Cell of nodeJ into cellJ, calcCellUe, and the exterior cells of
// cellUe by neumannCell (:226-297), in one launch.  An exterior cell's
// Neumann source is never exterior: its thread computes the source's cellJ
// and Ue as the source's own thread does (without counting the clamp twice).
template <typename Real>
__device__ __forceinline__ void node2cell_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& node,
                                             Real (&c)[3]) {
    const int m100 = nbr(g, i, j, k, -1, 0, 0), m110 = nbr(g, i, j, k, -1, -1, 0), m010 = nbr(g, i, j, k, 0, -1, 0);
    const int m001 = nbr(g, i, j, k, 0, 0, -1), m011 = nbr(g, i, j, k, 0, -1, -1), m101 = nbr(g, i, j, k, -1, 0, -1);
    const int m111 = nbr(g, i, j, k, -1, -1, -1);
    for (int l = 0; l < 3; ++l) {
        const Real* a = l == 0 ? node.x : (l == 1 ? node.y : node.z);
        c[l] = Real(0.125) * (a[n] + a[m100] + a[m110] + a[m010] + a[m001] + a[m011] + a[m101] + a[m111]);
    }
}

// node2Cell (:1092-1112): cell = mean of its 8 corner nodes
template <typename Real>
__global__ __launch_bounds__(256) void k_node2cell(Geom g, V3<Real> node, V3<Real> cell) {
    RH_CELL_PROLOGUE
    if (!all_nbrs(g, i, j, k)) return;
    Real c[3];
    node2cell_at(g, i, j, k, n, node, c);
    cell.x[n] = c[0]; cell.y[n] = c[1]; cell.z[n] = c[2];
}

template <typename Real>
__global__ __launch_bounds__(256) void k_cell_J_Ue(Geom g, V3<Real> nodeJ, V3<Real> cellJ, V3<Real> Ji,
                                                  Real* __restrict__ rho, const uint8_t* __restrict__ inner,
                                                  V3<Real> Ue, int use_hall, Real ox_, Real oy_, Real oz_,
                                                  Real max_ue2, Real min_rho, unsigned long long* counters) {
    RH_CELL_IJK
    bool c_ue = false, c_rho = false;
    if (in_grid) {
        if (all_nbrs(g, i, j, k)) {
            Real cj[3], u[3];
            node2cell_at(g, i, j, k, n, nodeJ, cj);
            cellJ.x[n] = cj[0]; cellJ.y[n] = cj[1]; cellJ.z[n] = cj[2];
            ue_at<Real, false>(n, cj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c_ue, c_rho);
            Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
        } else {
            const int src = neumann_source(g, i, j, k);
            if (src >= 0) {
                int si, sj, sk;
                cell_ijk(g, src, si, sj, sk);
                Real cj[3], u[3];
                bool c0 = false, c1 = false;
                node2cell_at(g, si, sj, sk, src, nodeJ, cj);
                ue_at<Real, false>(src, cj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c0, c1);
                Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
            }
        }
    }
    count_event(counters, C_MAX_CELL_UE, c_ue);
}

// upwindNodeB (:1268-1339): inverse-distance weights of the 8 cells around
// the node, seen from a point displaced by -Ue/|Ue|/2 (default (-1/2,0,0)).
template <typename Real>
__device__ __forceinline__ void upwind_B_at(const Geom& g, int i, int j, int k, const V3<Real>& cellB,
                                            const Real (&ue3)[3], Real (&b)[3]) {
    Real xu = Real(-0.5), yu = Real(0), zu = Real(0);
    const Real ue = sqrt(ue3[0] * ue3[0] + ue3[1] * ue3[1] + ue3[2] * ue3[2]);
    if (ue > Real(0)) {
        const Real a = Real(-0.5) / ue;
        xu = ue3[0] * a; yu = ue3[1] * a; zu = ue3[2] * a;
    }
    // reference order 111,112,121,211,122,221,212,222 (offsets 0 or +1)
    const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    Real w[8];
    int c[8];
    Real wsum = Real(0);
    for (int q = 0; q < 8; ++q) {
        const Real xc = off[q][0] ? Real(0.5) : Real(-0.5);
        const Real yc = off[q][1] ? Real(0.5) : Real(-0.5);
        const Real zc = off[q][2] ? Real(0.5) : Real(-0.5);
        w[q] = Real(1) / sqrt((xc - xu) * (xc - xu) + (yc - yu) * (yc - yu) + (zc - zu) * (zc - zu));
        c[q] = nbr(g, i, j, k, off[q][0], off[q][1], off[q][2]);
        wsum += w[q];
    }
    if (wsum > Real(0)) {
        Real bx = 0, by = 0, bz = 0;
        for (int q = 0; q < 8; ++q) {
            bx += w[q] * cellB.x[c[q]];
            by += w[q] * cellB.y[c[q]];
            bz += w[q] * cellB.z[c[q]];
        }
        b[0] = bx / wsum; b[1] = by / wsum; b[2] = bz / wsum;
    } else {
        b[0] = b[1] = b[2] = Real(0);
    }
}

// calcNodeE (:1391-1446): E = -Ue x B + eta J, |E| <= Emax
template <typename Real>
__device__ __forceinline__ bool node_E_at(const Real (&u)[3], const Real (&b)[3], Real eta, Real jx, Real jy,
                                          Real jz, Real max_E2, Real (&e)[3]) {
    Real ex = -(u[1] * b[2] - u[2] * b[1]);
    Real ey = -(u[2] * b[0] - u[0] * b[2]);
    Real ez = -(u[0] * b[1] - u[1] * b[0]);
    ex += eta * jx;
    ey += eta * jy;
    ez += eta * jz;
    bool c = false;
    if (max_E2 > Real(0)) {
        const Real e2 = ex * ex + ey * ey + ez * ez;
        if (e2 > max_E2) {
            const Real s = sqrt(max_E2 / e2);
            ex *= s; ey *= s; ez *= s;
            c = true;
        }
    }
    e[0] = ex; e[1] = ey; e[2] = ez;
    return c;
}

// calcNodeUe, upwindNodeB and calcNodeE (:300-352) in one launch: the three
// run on the same nodes (cells with their +x, +y, +z neighbours) and each
// reads only the node's own result of the one before (and cellB, nodeJ,
// nodeJi, nodeRhoQi, nodeEta, which none of them writes).
template <typename Real>
__global__ __launch_bounds__(256) void k_node_Ue_B_E(Geom g, V3<Real> J, V3<Real> Ji, Real* __restrict__ rho,
                                                    const uint8_t* __restrict__ inner, V3<Real> Ue, int use_hall,
                                                    Real ox_, Real oy_, Real oz_, Real max_ue2, Real min_rho,
                                                    V3<Real> cellB, V3<Real> nodeB, const Real* __restrict__ eta,
                                                    V3<Real> E, Real max_E2, unsigned long long* counters) {
    RH_CELL_IJK
    bool c_ue = false, c_rho = false, c_e = false;
    if (in_grid && !upper_missing(g, i, j, k)) {
        Real u[3], b[3], e[3];
        const Real jj[3] = {J.x[n], J.y[n], J.z[n]};
        ue_at<Real, true>(n, jj, Ji, rho, inner, use_hall, ox_, oy_, oz_, max_ue2, min_rho, u, c_ue, c_rho);
        Ue.x[n] = u[0]; Ue.y[n] = u[1]; Ue.z[n] = u[2];
        upwind_B_at(g, i, j, k, cellB, u, b);
        nodeB.x[n] = b[0]; nodeB.y[n] = b[1]; nodeB.z[n] = b[2];
        c_e = node_E_at(u, b, eta[n], J.x[n], J.y[n], J.z[n], max_E2, e);
        E.x[n] = e[0]; E.y[n] = e[1]; E.z[n] = e[2];
    }
    count_event(counters, C_MAX_NODE_UE, c_ue);
    count_event(counters, C_MIN_NODE_RHO, c_rho);
    count_event(counters, C_MAX_E, c_e);
}

// nodeAvg (:1115-1265).  The reference's gather buffer starts as
// {DBL_MAX, 0, 0, ...}: only a missing (-1,-1,-1) neighbour's x component is
// recognised as missing (weight 0 and renormalisation); other missing
// neighbours count as zeros with full weight.
template <typename Real>
__global__ __launch_bounds__(256) void k_node_avg(Geom g, V3<Real> Eold, V3<Real> E, Real c1, Real c2, Real c3, Real c4) {
    RH_CELL_PROLOGUE
    if (upper_missing(g, i, j, k)) return;
    const Real C[4] = {c1, c2, c3, c4};
    // reference summation order of the 27 nodes (nodeEl111, 110, 011, ...)
    const int8_t ord[27][3] = {{0, 0, 0}, {0, 0, -1}, {-1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, -1, 0}, {0, 0, 1},
                               {-1, 0, -1}, {0, -1, -1}, {0, 1, -1}, {1, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
                               {1, -1, 0}, {-1, 0, 1}, {0, 1, 1}, {1, 0, 1}, {0, -1, 1}, {-1, -1, 1}, {-1, 1, 1},
                               {1, -1, 1}, {1, 1, 1}, {-1, -1, -1}, {-1, 1, -1}, {1, -1, -1}, {1, 1, -1}};
    for (int l = 0; l < 3; ++l) {
        const Real* a = l == 0 ? Eold.x : (l == 1 ? Eold.y : Eold.z);
        Real w[27], v[27];
        bool zero_found = false;
        for (int q = 0; q < 27; ++q) {
            const int m = nbr(g, i, j, k, ord[q][0], ord[q][1], ord[q][2]);
            const int dist = (ord[q][0] != 0) + (ord[q][1] != 0) + (ord[q][2] != 0);
            w[q] = C[dist];
            if (m >= 0) v[q] = a[m];
            else v[q] = Real(0);
            if (m < 0 && l == 0 && q == 23) { w[q] = Real(0); zero_found = true; }
        }
        if (zero_found) {
            Real cs = Real(0);
            for (int q = 0; q < 27; ++q) cs += w[q];
            if (cs > Real(0))
                for (int q = 0; q < 27; ++q) w[q] /= cs;
        }
        Real acc = Real(0);
        for (int q = 0; q < 27; ++q) acc += w[q] * v[q];
        (l == 0 ? E.x : (l == 1 ? E.y : E.z))[n] = acc;
    }
}

// faceCurl, Faraday (:1587-1671): faceB -= dt curl(nodeE)
template <typename Real>
__device__ __forceinline__ void face_curl_at(const Geom& g, int i, int j, int k, int n, const V3<Real>& E,
                                             const V3<Real>& F, Real dt) {
    const bool xn = has_neg(i, g.per[0]), yn = has_neg(j, g.per[1]), zn = has_neg(k, g.per[2]);
    const bool doX = yn && zn, doY = xn && zn, doZ = xn && yn;
    const Real dx = (Real)g.dx;
    // node1 (-1,-1,0), node3 (-1,0,-1), node4 (-1,0,0), node5 (0,-1,0),
    // node6 (0,-1,-1), node7 (0,0,-1), node8 (0,0,0)
    if (doX) {
        const int n5 = nbr(g, i, j, k, 0, -1, 0), n6 = nbr(g, i, j, k, 0, -1, -1), n7 = nbr(g, i, j, k, 0, 0, -1);
        const Real curl = Real(0.5) * (+E.z[n5] + E.z[n6] - E.y[n6] - E.y[n7] - E.z[n7] - E.z[n] + E.y[n] + E.y[n5]) / dx;
        F.x[n] += dt * curl;
    }
    if (doY) {
        const int n3 = nbr(g, i, j, k, -1, 0, -1), n4 = nbr(g, i, j, k, -1, 0, 0), n7 = nbr(g, i, j, k, 0, 0, -1);
        const Real curl = Real(0.5) * (-E.x[n4] - E.x[n] + E.z[n] + E.z[n7] + E.x[n7] + E.x[n3] - E.z[n3] - E.z[n4]) / dx;
        F.y[n] += dt * curl;
    }
    if (doZ) {
        const int n1 = nbr(g, i, j, k, -1, -1, 0), n4 = nbr(g, i, j, k, -1, 0, 0), n5 = nbr(g, i, j, k, 0, -1, 0);
        const Real curl = Real(0.5) * (-E.x[n1] - E.x[n5] - E.y[n5] - E.y[n] + E.x[n] + E.x[n4] + E.y[n4] + E.y[n1]) / dx;
        F.z[n] += dt * curl;
    }
}

// calcCellEp (:1674-1737): Ep = -coeff grad(rho)/rho (isothermal) or -coeff grad(rho)
template <typename Real>
__device__ __forceinline__ void cell_Ep_at(const Geom& g, int i, int j, int k, int n, const Real* __restrict__ nrho,
                                           const Real* __restrict__ crho, const uint8_t* __restrict__ inner,
                                           const V3<Real>& Ep, Real coeff, int isothermal) {
    Ep.x[n] = Ep.y[n] = Ep.z[n] = Real(0);
    if (inner[n]) return;
    // node values at offsets (a,b,c) in {-1,0}^3
    auto N = [&](int a, int b, int c) { return nrho[nbr(g, i, j, k, a, b, c)]; };
    const Real xP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(0, -1, -1) + N(0, -1, 0));
    const Real xN = Real(0.25) * (N(-1, 0, 0) + N(-1, 0, -1) + N(-1, -1, -1) + N(-1, -1, 0));
    const Real yP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(-1, 0, -1) + N(-1, 0, 0));
    const Real yN = Real(0.25) * (N(0, -1, 0) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, -1, 0));
    const Real zP = Real(0.25) * (N(0, 0, 0) + N(0, -1, 0) + N(-1, -1, 0) + N(-1, 0, 0));
    const Real zN = Real(0.25) * (N(0, 0, -1) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, 0, -1));
    const Real dx = (Real)g.dx;
    const Real gx = (xP - xN) / dx, gy = (yP - yN) / dx, gz = (zP - zN) / dx;
    if (!isothermal) {
        Ep.x[n] = -coeff * gx; Ep.y[n] = -coeff * gy; Ep.z[n] = -coeff * gz;
    } else if (fabs(crho[n]) > Real(0)) {
        Ep.x[n] = -coeff * gx / crho[n]; Ep.y[n] = -coeff * gy / crho[n]; Ep.z[n] = -coeff * gz / crho[n];
    }
}

// faceCurl (Faraday) and calcCellEp (:406-432), then, when out is not null,
// the cell's entries of the per-tile push-field blocks (k_push_fields) from
// its final faceB, cellUe and cellEp, in one launch: each stage reads only
// arrays no stage writes (nodeE, node and cell rho, flags, cellUe) and the
// cell's own results.
template <typename Real>
__global__ __launch_bounds__(256) void k_face_curl_Ep(Geom g, V3<Real> E, V3<Real> F, Real dt, int use_e_pressure,
                                                     const Real* __restrict__ nrho, const Real* __restrict__ crho,
                                                     const uint8_t* __restrict__ inner, V3<Real> Ep, Real coeff,
                                                     int isothermal, V3<Real> Ue, uint32_t* __restrict__ out) {
    RH_CELL_PROLOGUE
    if (!upper_missing(g, i, j, k)) face_curl_at(g, i, j, k, n, E, F, dt);
    if (use_e_pressure && all_nbrs(g, i, j, k)) cell_Ep_at(g, i, j, k, n, nrho, crho, inner, Ep, coeff, isothermal);
    if (out) {
        const PushFields<Real> pf{F.x, F.y, F.z, Ue.x, Ue.y, Ue.z, Ep.x, Ep.y, Ep.z};
        push_fields_at(g, i, j, k, n, pf, use_e_pressure, out);
    }
}

// The three inner flags the fused field update reads, one byte per cell
// (bit 0 innerFlagField, bit 1 innerFlagNode, bit 2 innerFlagCellEp),
// rebuilt whenever one of them is uploaded.
__global__ __launch_bounds__(256) void k_pack_inner_flags(int nc, const uint8_t* __restrict__ field,
                                                         const uint8_t* __restrict__ node,
                                                         const uint8_t* __restrict__ cellEp, uint8_t* __restrict__ out) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n < nc) out[n] = (field[n] ? 1 : 0) | (node[n] ? 2 : 0) | (cellEp[n] ? 4 : 0);
}

// ---- the fused field update ------------------------------------------------
//
// propagateB without E smoothing (src/hybrid_propagator.cpp:42-440, steps
// 1-8 and 14-15) in one launch over field tiles of FT cells, with every
// stencil input staged in LDS:
//   faceB of the tile grown by two cells below and one above (after the
//   exterior cells' neumannFace and setIMFFace copies), cellRhoQi, cellJi
//   and cellB of the tile grown by one cell (exterior cells: neumannCell
//   copies of their source and setIMFCell; interior cellB by face2Cell),
//   then the nodes of the tile grown by one node below (cell2Node of cellB,
//   cellRhoQi and cellJi, calcNodeJ, calcNodeUe, upwindNodeB, calcNodeE) in
//   LDS, then for the tile's own cells node2Cell of nodeJ with calcCellUe
//   (exterior cells: their source's value), calcCellEp and faceCurl.  A
//   node of the halo is computed by every tile that holds it, and counted
//   only by the tile that owns it.  faceB is read from one buffer and
//   written to another (tiles read their neighbours' old faces as halo).
//   The arithmetic of every stage is the same as in the kernels above.
// The exterior cells' Neumann sources lie in the same tile when no tile is
// one cell thick along a non-periodic axis (the host checks n % FT != 1).
template <typename Real> struct FieldTile;
template <> struct FieldTile<float> { static constexpr int X = 8, Y = 8, Z = 8; };
template <> struct FieldTile<double> { static constexpr int X = 8, Y = 4, Z = 4; };

template <typename Real>
struct FieldUpdateArgs {
    V3<Real> faceB_in, faceB_out;
    Real* rho;                // cellRhoQi (exterior cells: their Neumann copies, full update)
    V3<Real> Ji;              // cellJi (idem)
    const Real* eta;          // nodeEta
    const uint8_t* inner3;    // per cell: bit 0 innerFlagField, bit 1 innerFlagNode, bit 2 innerFlagCellEp
    V3<Real> cellUe, cellEp;
    // the reference's other arrays (full update; cellB also when write_cellB)
    V3<Real> cellB, cellJ, nodeB, nodeJi, nodeJ, nodeUe, nodeE;
    Real* nodeRho;
    uint32_t* push_blocks;    // the next push's per-tile field blocks (null: none)
    double* aveB;             // cellAverageB += cellB (temporal averages; null: off)
    unsigned long long* counters;
    Real dt, e_coeff, max_ue2, min_rho, max_vw, max_E2;
    Real oue[3];              // fieldObstacleUe
    Real imf[3];
    int imf_cell, imf_face;
    int use_hall, use_e_pressure, isothermal;
    int full, write_cellB;
};

// 1/sqrt(x): the hardware reciprocal square root in FP32 (1 ulp), the
// correctly rounded quotient in FP64
__device__ __forceinline__ float inv_sqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double inv_sqrt(double x) { return 1.0 / sqrt(x); }
// 1/x and sqrt(x): the hardware instructions in FP32 (1 ulp), correctly
// rounded in FP64
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }

// calcCellUe / calcNodeUe on values (ue_at above, NODE clamps r)
template <typename Real, bool NODE>
__device__ __forceinline__ void ue_core(const Real (&jj)[3], const Real (&ji)[3], Real& r, bool inner, int use_hall,
                                        const Real (&oue)[3], Real max_ue2, Real min_rho, Real (&u)[3], bool& c_ue,
                                        bool& c_rho) {
    if (inner) {
        u[0] = oue[0]; u[1] = oue[1]; u[2] = oue[2];
        return;
    }
    if (NODE && r < min_rho) { r = min_rho; c_rho = true; }
    const Real ir = fabs(r) > Real(0) ? fast_rcp(r) : Real(0);
    for (int l = 0; l < 3; ++l) u[l] = use_hall ? -(jj[l] - ji[l]) * ir : ji[l] * ir;
    const Real u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (u2 > max_ue2) {
        const Real norm = sqrt(max_ue2 / u2);
        u[0] *= norm; u[1] *= norm; u[2] *= norm;
        c_ue = true;
    }
}

// setIMFCell / setIMFFace on a cell's values (imf_at above)
template <typename Real>
__device__ __forceinline__ void imf_values(const Geom& g, int i, int j, int k, int face, int flags, const Real (&b)[3],
                                           Real (&v)[3]) {
    const int miss = missing_mask(g, i, j, k);
    const int wall_bit[6] = {2, 1, 8, 4, 32, 16};
    for (int w = 0; w < 6; ++w) {
        if (!((flags >> w) & 1) || !(miss & wall_bit[w])) continue;
        const int axis = w / 2;
        if (!face || axis != 0) v[0] = b[0];
        if (!face || axis != 1) v[1] = b[1];
        if (!face || axis != 2) v[2] = b[2];
    }
}

// Threads of the fused update: a tile's loops (1331 halo faces, 1000 cells,
// 729 nodes, 512 own cells for FP32 8x8x8 tiles) take one or two rounds each
// (768 threads: the node stage in one round; 133 us per update on the
// 128x96x96 box against 151 us with 512 threads and 180 us with 8x8x4
// tiles of 256 threads).
template <typename Real> struct FieldUpdateThreads;
template <> struct FieldUpdateThreads<float> { static constexpr int N = 768; };
template <> struct FieldUpdateThreads<double> { static constexpr int N = 256; };

template <typename Real, bool FULL>
__global__ __launch_bounds__(FieldUpdateThreads<Real>::N) void k_field_update(Geom g, FieldUpdateArgs<Real> a) {
    constexpr int TX = FieldTile<Real>::X, TY = FieldTile<Real>::Y, TZ = FieldTile<Real>::Z;
    constexpr int X2 = TX + 3, Y2 = TY + 3, Z2 = TZ + 3, N2 = X2 * Y2 * Z2;  // faceB: cells -2..T
    constexpr int X1 = TX + 2, Y1 = TY + 2, Z1 = TZ + 2, N1 = X1 * Y1 * Z1;  // cells -1..T
    constexpr int XN = TX + 1, YN = TY + 1, ZN = TZ + 1, NN = XN * YN * ZN;  // nodes -1..T-1
    constexpr int NT = FieldUpdateThreads<Real>::N;
    constexpr int R2 = (N2 + NT - 1) / NT;  // rounds over the face halo
    __shared__ Real sF[3][N2];
    __shared__ Real sB[3][N1], sJi[3][N1], sR[N1];
    __shared__ Real nJ[3][NN], nE[3][NN], nR[NN];
    // per cell of the cell region: nodeEta of its node, and the inner flags
    // (bit 0 innerFlagField, bit 1 innerFlagNode, bit 2 innerFlagCellEp)
    __shared__ Real sEta[N1];
    __shared__ uint8_t sFl[N1];
    const int tid = threadIdx.x;
    // tiles x-fastest, dealt to the XCDs in contiguous ranges: neighbouring
    // tiles, which read the same halo lines, share an L2 and run close in time
    const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY, ntz = (g.nz + TZ - 1) / TZ;
    const int T = xcd_index(ntx * nty * ntz, (int)blockIdx.x);
    if (T >= ntx * nty * ntz) return;
    const int ox = (T % ntx) * TX, oy = ((T / ntx) % nty) * TY, oz = (T / (ntx * nty)) * TZ;
    const V3<Real>& F = a.faceB_in;
    // a tile whose halo lies inside the grid without a wrap and holds no cell
    // of the grid's outer layer: no wrap, no wall rule (the same for the
    // whole workgroup)
    const bool deep = ox - 2 >= 1 && oy - 2 >= 1 && oz - 2 >= 1 && ox + TX <= g.nx - 2 && oy + TY <= g.ny - 2 &&
                      oz + TZ <= g.nz - 2;

    // 1. faceB, cellRhoQi, cellJi, nodeEta and the inner flags of the halo.
    //    All of a thread's loads are issued before its LDS stores; the cells
    //    without all neighbours then get their copies (neumannFace and
    //    setIMFFace on faceB; the Neumann source's face2Cell with setIMFCell
    //    for cellB, and its cellRhoQi, cellJi).
    int en[R2];                       // cell index of the entry (-1: none or outside the grid)
    int ei[R2], ej[R2], ek[R2];       // its grid coordinates
    Real lf[R2][3], lj[R2][3], lr[R2], le[R2];
    uint8_t lfl[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int e = tid + r * NT;
        en[r] = -1;
        lf[r][0] = lf[r][1] = lf[r][2] = Real(0);
        lj[r][0] = lj[r][1] = lj[r][2] = Real(0);
        lr[r] = le[r] = Real(0);
        lfl[r] = 0;
        if (e >= N2) continue;
        const int lx = e % X2, ly = (e / X2) % Y2, lz = e / (X2 * Y2);
        int i = ox - 2 + lx, j = oy - 2 + ly, k = oz - 2 + lz;
        if (!deep) {
            i = wrap_axis(i, g.nx, g.per[0]); j = wrap_axis(j, g.ny, g.per[1]); k = wrap_axis(k, g.nz, g.per[2]);
            if (i < 0 || j < 0 || k < 0) continue;
        }
        const int n = cell_index(g, i, j, k);
        en[r] = n;
        ei[r] = i; ej[r] = j; ek[r] = k;
        lf[r][0] = F.x[n]; lf[r][1] = F.y[n]; lf[r][2] = F.z[n];
        if (lx >= 1 && ly >= 1 && lz >= 1) {
            lj[r][0] = a.Ji.x[n]; lj[r][1] = a.Ji.y[n]; lj[r][2] = a.Ji.z[n];
            lr[r] = a.rho[n];
            le[r] = a.eta[n];
            lfl[r] = a.inner3[n];
        }
    }
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int e = tid + r * NT;
        if (e >= N2) continue;
        const int lx = e % X2, ly = (e / X2) % Y2, lz = e / (X2 * Y2);
        const bool in1 = lx >= 1 && ly >= 1 && lz >= 1;
        const int e1 = in1 ? (lx - 1) + X1 * ((ly - 1) + Y1 * (lz - 1)) : 0;
        sF[0][e] = lf[r][0]; sF[1][e] = lf[r][1]; sF[2][e] = lf[r][2];
        if (in1) {
            sB[0][e1] = sB[1][e1] = sB[2][e1] = Real(0);
            sJi[0][e1] = lj[r][0]; sJi[1][e1] = lj[r][1]; sJi[2][e1] = lj[r][2];
            sR[e1] = lr[r];
            sEta[e1] = le[r];
            sFl[e1] = lfl[r];
        }
        const int n = en[r];
        if (n < 0) continue;
        const int i = ei[r], j = ej[r], k = ek[r];
        if (deep || all_nbrs(g, i, j, k)) continue;
        const int miss = missing_mask(g, i, j, k);
        Real f[3];
        for (int l = 0; l < 3; ++l) {
            const FaceSrc R = c_face_src.t[miss][l];
            const int src = R.found ? nbr(g, i, j, k, R.a, R.b, R.c) : -1;
            f[l] = F.c(l, src >= 0 ? src : n);
        }
        if (a.imf_face) imf_values(g, i, j, k, 1, a.imf_face, a.imf, f);
        sF[0][e] = f[0]; sF[1][e] = f[1]; sF[2][e] = f[2];
        if (in1) {
            const int src = neumann_source(g, i, j, k);
            int si, sj, sk;
            cell_ijk(g, src, si, sj, sk);
            Real c[3];
            face2cell_at(g, si, sj, sk, src, F, c[0], c[1], c[2]);
            if (a.imf_cell) imf_values(g, i, j, k, 0, a.imf_cell, a.imf, c);
            sB[0][e1] = c[0]; sB[1][e1] = c[1]; sB[2][e1] = c[2];
            sJi[0][e1] = a.Ji.x[src]; sJi[1][e1] = a.Ji.y[src]; sJi[2][e1] = a.Ji.z[src];
            sR[e1] = a.rho[src];
        }
    }
    __syncthreads();
    // 2. face2Cell of the interior cells (their lower faces are never
    //    changed by the exterior copies: see k_cell_B_bc)
    for (int e = tid; e < N1; e += blockDim.x) {
        const int lx = e % X1, ly = (e / X1) % Y1, lz = e / (X1 * Y1);
        if (!deep) {
            const int i = wrap_axis(ox - 1 + lx, g.nx, g.per[0]), j = wrap_axis(oy - 1 + ly, g.ny, g.per[1]),
                      k = wrap_axis(oz - 1 + lz, g.nz, g.per[2]);
            if (i < 0 || j < 0 || k < 0 || !all_nbrs(g, i, j, k)) continue;
        }
        const int f = (lx + 1) + X2 * ((ly + 1) + Y2 * (lz + 1));
        sB[0][e] = Real(0.5) * (sF[0][f] + sF[0][f - 1]);
        sB[1][e] = Real(0.5) * (sF[1][f] + sF[1][f - X2]);
        sB[2][e] = Real(0.5) * (sF[2][f] + sF[2][f - X2 * Y2]);
    }
    __syncthreads();
    // 3. the nodes: cell2Node, calcNodeJ, calcNodeUe, upwindNodeB, calcNodeE
    for (int q = tid; q < NN; q += blockDim.x) {
        const int qx = q % XN, qy = (q / XN) % YN, qz = q / (XN * YN);
        int i = ox - 1 + qx, j = oy - 1 + qy, k = oz - 1 + qz;
        bool ok = true;
        if (!deep) {
            i = wrap_axis(i, g.nx, g.per[0]); j = wrap_axis(j, g.ny, g.per[1]); k = wrap_axis(k, g.nz, g.per[2]);
            ok = i >= 0 && j >= 0 && k >= 0 && !upper_missing(g, i, j, k);
        }
        // the node of one of the tile's own cells (a partial tile's halo may
        // wrap onto cells of another tile along a periodic axis)
        const bool own = ok && qx >= 1 && qy >= 1 && qz >= 1 && ox - 1 + qx < g.nx && oy - 1 + qy < g.ny &&
                         oz - 1 + qz < g.nz;
        bool c_vw = false, c_ue = false, c_rho = false, c_e = false;
        if (ok) {
            const int n = cell_index(g, i, j, k);
            const int c = qx + X1 * (qy + Y1 * qz);  // the node's cell in the cell region
            const int c001 = c + X1 * Y1, c010 = c + X1, c100 = c + 1;
            const int c011 = c010 + X1 * Y1, c110 = c010 + 1, c101 = c001 + 1, c111 = c011 + 1;
            auto avg = [&](const Real* v) {
                return Real(0.125) * (v[c] + v[c001] + v[c010] + v[c100] + v[c011] + v[c110] + v[c101] + v[c111]);
            };
            const Real bx = avg(sB[0]), by = avg(sB[1]), bz = avg(sB[2]);
            Real r = avg(sR);
            const Real ji[3] = {avg(sJi[0]), avg(sJi[1]), avg(sJi[2])};
            // calcNodeJ from faceB of the cells n, +x, +y, +z, +xy, +xz, +yz
            const int f = (qx + 1) + X2 * ((qy + 1) + Y2 * (qz + 1));
            const int f100 = f + 1, f010 = f + X2, f001 = f + X2 * Y2;
            const int f110 = f010 + 1, f101 = f001 + 1, f011 = f001 + X2;
            const Real ex1 = -sF[2][f] + sF[1][f] + sF[2][f010] - sF[1][f001];
            const Real ex2 = -sF[2][f100] + sF[1][f100] + sF[2][f110] - sF[1][f101];
            const Real ey1 = +sF[2][f] + sF[0][f001] - sF[2][f100] - sF[0][f];
            const Real ey2 = +sF[2][f010] + sF[0][f011] - sF[2][f110] - sF[0][f010];
            const Real ez1 = -sF[1][f] + sF[0][f] + sF[1][f100] - sF[0][f010];
            const Real ez2 = -sF[1][f001] + sF[0][f001] + sF[1][f101] - sF[0][f011];
            const Real qe = Real(1.60217653e-19);
            const Real mu0 = Real(4.0e-7 * 3.14159265358979323846);
            Real d = Real(1), vw = Real(0);
            const Real ne = r * Real(1.0 / 1.60217653e-19);
            const Real Btot = fast_sqrt(bx * bx + by * by + bz * bz);
            const Real dx = (Real)g.dx;
            if (ne > Real(0) && dx > Real(0))
                vw = Real(2) * Btot * Real(3.14159265358979323846) * fast_rcp(mu0 * ne * qe * dx);
            if (vw > a.max_vw && a.max_vw > Real(0)) { d = vw / a.max_vw; c_vw = true; }
            const Real af = Real(0.5) * fast_rcp(dx * mu0 * d);
            const Real J[3] = {(ex1 + ex2) * af, (ey1 + ey2) * af, (ez1 + ez2) * af};
            // calcNodeUe (clamps the node's rho)
            Real u[3];
            ue_core<Real, true>(J, ji, r, (sFl[c] & 2) != 0, a.use_hall, a.oue, a.max_ue2, a.min_rho, u, c_ue, c_rho);
            // upwindNodeB: cells in the reference's order 111,112,121,211,122,221,212,222
            Real xu = Real(-0.5), yu = Real(0), zu = Real(0);
            const Real ue = fast_sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (ue > Real(0)) {
                const Real am = Real(-0.5) * fast_rcp(ue);
                xu = u[0] * am; yu = u[1] * am; zu = u[2] * am;
            }
            const int cc[8] = {c, c001, c010, c100, c011, c110, c101, c111};
            const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
            Real w[8], wsum = Real(0);
            for (int t = 0; t < 8; ++t) {
                const Real xc = off[t][0] ? Real(0.5) : Real(-0.5);
                const Real yc = off[t][1] ? Real(0.5) : Real(-0.5);
                const Real zc = off[t][2] ? Real(0.5) : Real(-0.5);
                w[t] = inv_sqrt((xc - xu) * (xc - xu) + (yc - yu) * (yc - yu) + (zc - zu) * (zc - zu));
                wsum += w[t];
            }
            Real b[3];
            if (wsum > Real(0)) {
                Real b0 = 0, b1 = 0, b2 = 0;
                for (int t = 0; t < 8; ++t) {
                    b0 += w[t] * sB[0][cc[t]];
                    b1 += w[t] * sB[1][cc[t]];
                    b2 += w[t] * sB[2][cc[t]];
                }
                const Real iw = fast_rcp(wsum);
                b[0] = b0 * iw; b[1] = b1 * iw; b[2] = b2 * iw;
            } else {
                b[0] = b[1] = b[2] = Real(0);
            }
            Real E[3];
            c_e = node_E_at(u, b, sEta[c], J[0], J[1], J[2], a.max_E2, E);
            for (int l = 0; l < 3; ++l) { nJ[l][q] = J[l]; nE[l][q] = E[l]; }
            nR[q] = r;
            if (FULL && own) {
                a.nodeB.x[n] = b[0]; a.nodeB.y[n] = b[1]; a.nodeB.z[n] = b[2];
                a.nodeRho[n] = r;
                a.nodeJi.x[n] = ji[0]; a.nodeJi.y[n] = ji[1]; a.nodeJi.z[n] = ji[2];
                a.nodeJ.x[n] = J[0]; a.nodeJ.y[n] = J[1]; a.nodeJ.z[n] = J[2];
                a.nodeUe.x[n] = u[0]; a.nodeUe.y[n] = u[1]; a.nodeUe.z[n] = u[2];
                a.nodeE.x[n] = E[0]; a.nodeE.y[n] = E[1]; a.nodeE.z[n] = E[2];
            }
        }
        count_event(a.counters, C_MAX_VW, c_vw && own);
        count_event(a.counters, C_MAX_NODE_UE, c_ue && own);
        count_event(a.counters, C_MIN_NODE_RHO, c_rho && own);
        count_event(a.counters, C_MAX_E, c_e && own);
    }
    __syncthreads();
    // 4. the tile's cells: node2Cell + calcCellUe (exterior cells: the
    //    source's value), calcCellEp, faceCurl, the push-field blocks
    for (int t = tid; t < TX * TY * TZ; t += blockDim.x) {
        const int rx = t % TX, ry = (t / TX) % TY, rz = t / (TX * TY);
        const int i = ox + rx, j = oy + ry, k = oz + rz;
        bool c_ue = false;
        if (i < g.nx && j < g.ny && k < g.nz) {
            const int n = cell_index(g, i, j, k);
            const int c1 = (rx + 1) + X1 * ((ry + 1) + Y1 * (rz + 1));
            const bool interior = deep || all_nbrs(g, i, j, k);
            // node2Cell of nodeJ at the cell whose node sits at node index qn
            auto cellJ_at = [&](int qn, Real (&cj)[3]) {
                const int m100 = qn - 1, m010 = qn - XN, m001 = qn - XN * YN;
                const int m110 = m010 - 1, m011 = m001 - XN, m101 = m001 - 1, m111 = m011 - 1;
                for (int l = 0; l < 3; ++l) {
                    const Real* v = nJ[l];
                    cj[l] = Real(0.125) * (v[qn] + v[m100] + v[m110] + v[m010] + v[m001] + v[m011] + v[m101] + v[m111]);
                }
            };
            const int qn = (rx + 1) + XN * ((ry + 1) + YN * (rz + 1));
            Real u[3], cj[3] = {Real(0), Real(0), Real(0)};
            bool c_rho = false;
            if (interior) {
                cellJ_at(qn, cj);
                Real r = sR[c1];
                const Real ji[3] = {sJi[0][c1], sJi[1][c1], sJi[2][c1]};
                ue_core<Real, false>(cj, ji, r, (sFl[c1] & 1) != 0, a.use_hall, a.oue, a.max_ue2, a.min_rho, u, c_ue,
                                     c_rho);
            } else {
                const int src = neumann_source(g, i, j, k);
                int si, sj, sk;
                cell_ijk(g, src, si, sj, sk);
                const int dxs = si - i, dys = sj - j, dzs = sk - k;  // -1, 0 or +1 (no wrap: exterior)
                Real cs[3];
                cellJ_at(qn + dxs + XN * (dys + YN * dzs), cs);
                const int cs1 = c1 + dxs + X1 * (dys + Y1 * dzs);
                Real r = sR[cs1];
                const Real ji[3] = {sJi[0][cs1], sJi[1][cs1], sJi[2][cs1]};
                bool c0 = false;
                ue_core<Real, false>(cs, ji, r, (sFl[cs1] & 1) != 0, a.use_hall, a.oue, a.max_ue2, a.min_rho, u, c0,
                                     c_rho);
            }
            a.cellUe.x[n] = u[0]; a.cellUe.y[n] = u[1]; a.cellUe.z[n] = u[2];
            // calcCellEp from the clamped nodeRhoQi
            Real ep[3] = {Real(0), Real(0), Real(0)};
            if (a.use_e_pressure) {
                if (interior) {
                    if (!(sFl[c1] & 4)) {
                        auto N = [&](int ax, int ay, int az) { return nR[qn + ax + XN * (ay + YN * az)]; };
                        const Real xP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(0, -1, -1) + N(0, -1, 0));
                        const Real xN = Real(0.25) * (N(-1, 0, 0) + N(-1, 0, -1) + N(-1, -1, -1) + N(-1, -1, 0));
                        const Real yP = Real(0.25) * (N(0, 0, 0) + N(0, 0, -1) + N(-1, 0, -1) + N(-1, 0, 0));
                        const Real yN = Real(0.25) * (N(0, -1, 0) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, -1, 0));
                        const Real zP = Real(0.25) * (N(0, 0, 0) + N(0, -1, 0) + N(-1, -1, 0) + N(-1, 0, 0));
                        const Real zN = Real(0.25) * (N(0, 0, -1) + N(0, -1, -1) + N(-1, -1, -1) + N(-1, 0, -1));
                        const Real idx = Real(1) / (Real)g.dx;
                        const Real gx = (xP - xN) * idx, gy = (yP - yN) * idx, gz = (zP - zN) * idx;
                        const Real crho = sR[c1];
                        if (!a.isothermal) {
                            ep[0] = -a.e_coeff * gx; ep[1] = -a.e_coeff * gy; ep[2] = -a.e_coeff * gz;
                        } else if (fabs(crho) > Real(0)) {
                            const Real ic = fast_rcp(crho);
                            ep[0] = -a.e_coeff * gx * ic; ep[1] = -a.e_coeff * gy * ic; ep[2] = -a.e_coeff * gz * ic;
                        }
                    }
                    a.cellEp.x[n] = ep[0]; a.cellEp.y[n] = ep[1]; a.cellEp.z[n] = ep[2];
                } else {
                    // cells the reference's calcCellEp skips keep their values
                    ep[0] = a.cellEp.x[n]; ep[1] = a.cellEp.y[n]; ep[2] = a.cellEp.z[n];
                }
            }
            // faceCurl (Faraday) on the faces after their exterior copies
            const int f = (rx + 2) + X2 * ((ry + 2) + Y2 * (rz + 2));
            Real fb[3] = {sF[0][f], sF[1][f], sF[2][f]};
            if (deep || !upper_missing(g, i, j, k)) {
                const bool xn = deep || has_neg(i, g.per[0]), yn = deep || has_neg(j, g.per[1]),
                           zn = deep || has_neg(k, g.per[2]);
                const Real idx = Real(1) / (Real)g.dx;
                auto E = [&](int l, int ax, int ay, int az) { return nE[l][qn + ax + XN * (ay + YN * az)]; };
                if (yn && zn) {
                    const Real curl = Real(0.5) * (+E(2, 0, -1, 0) + E(2, 0, -1, -1) - E(1, 0, -1, -1) - E(1, 0, 0, -1) -
                                                   E(2, 0, 0, -1) - E(2, 0, 0, 0) + E(1, 0, 0, 0) + E(1, 0, -1, 0)) * idx;
                    fb[0] += a.dt * curl;
                }
                if (xn && zn) {
                    const Real curl = Real(0.5) * (-E(0, -1, 0, 0) - E(0, 0, 0, 0) + E(2, 0, 0, 0) + E(2, 0, 0, -1) +
                                                   E(0, 0, 0, -1) + E(0, -1, 0, -1) - E(2, -1, 0, -1) - E(2, -1, 0, 0)) * idx;
                    fb[1] += a.dt * curl;
                }
                if (xn && yn) {
                    const Real curl = Real(0.5) * (-E(0, -1, -1, 0) - E(0, 0, -1, 0) - E(1, 0, -1, 0) - E(1, 0, 0, 0) +
                                                   E(0, 0, 0, 0) + E(0, -1, 0, 0) + E(1, -1, 0, 0) + E(1, -1, -1, 0)) * idx;
                    fb[2] += a.dt * curl;
                }
            }
            a.faceB_out.x[n] = fb[0]; a.faceB_out.y[n] = fb[1]; a.faceB_out.z[n] = fb[2];
            if (FULL || a.write_cellB) {
                a.cellB.x[n] = sB[0][c1]; a.cellB.y[n] = sB[1][c1]; a.cellB.z[n] = sB[2][c1];
            }
            if (a.aveB) {
                a.aveB[n] += (double)sB[0][c1];
                a.aveB[(size_t)g.nc + n] += (double)sB[1][c1];
                a.aveB[2 * (size_t)g.nc + n] += (double)sB[2][c1];
            }
            if (FULL) {
                if (interior) {
                    a.cellJ.x[n] = cj[0]; a.cellJ.y[n] = cj[1]; a.cellJ.z[n] = cj[2];
                } else {
                    a.rho[n] = sR[c1];
                    a.Ji.x[n] = sJi[0][c1]; a.Ji.y[n] = sJi[1][c1]; a.Ji.z[n] = sJi[2][c1];
                }
            }
            if (a.push_blocks)
                push_fields_store(g, i, j, k, fb[0], fb[1], fb[2], u[0], u[1], u[2], ep[0], ep[1], ep[2], a.use_e_pressure,
                                  a.push_blocks, false, deep);
        }
        count_event(a.counters, C_MAX_CELL_UE, c_ue);
    }
}

}  // namespace rh
