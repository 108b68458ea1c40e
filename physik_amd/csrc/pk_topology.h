// Host-side topology of the implicit system: everything that depends only on
// connectivity, fixed/active flags and material, built once per change and
// uploaded. Deterministic integer work (the BSR pattern and the element->slot
// gather lists are bit-exact functions of the inputs).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "physik_amd.h"

namespace pk
{

// Per-tet parameter block, stored structure-of-arrays: prm[k * T + t].
enum TetParam
{
    P_DMINV = 0,    // 9 values, Dm^-1 row-major (rows are the shape gradients g1..g3)
    P_EDGE = 9,     // 9 values, rest edges X_a - X_0 for a = 1..3 (xyz each)
    P_VOL = 18,
    P_LAMBDA = 19,
    P_MU = 20,
    P_DAMP = 21,
    P_GDOT = 22,    // 10 values g_a . g_b for the local pairs a <= b (written on the device)
    P_COUNT = 32
};

// Element output of one substep:
//   basis   one 144-byte record per tet, basis[t * EB_COUNT + k]: R g_a
//           (k = 3a + i), then R R^T (xx, xy, xz, yy, yz, zz). The blocks
//           R K_ab R^T are not stored; the slot gather rebuilds each from
//           these with the closed form the element kernel also uses, and
//           finds all of a contribution's basis values in the tet's one or
//           two cache lines.
//   corner  3 doubles per (dynamic row, tet, corner) in the order of the row
//           lists, written at cpos[a * T + t] so that the right-hand side
//           reads a contiguous run: implicit, dt f_a - dt^2 sum_b (R K_ab R^T)
//           v_b (the force dropped when not finite); explicit, f_a.
// Reverse mode keeps its corner terms in the order of the node gather lists
// (alist), N_ADJ doubles per (node, tet, corner) written at apos[a * T + t]:
// dPhi/dx_a (3), then dPhi/dv_a (3), so that a node reads one contiguous run.
enum ElemLayout
{
    EB_RG = 0,
    EB_RRT = 12,
    EB_COUNT = 18,
    N_CORNER = 3,  // doubles per corner entry
    N_ADJ = 6      // doubles per reverse-mode corner entry
};

// Constant stiffness factors of one tet, one 96-byte record per tet
// kdat[t * KD_COUNT + k] (built on the device from the parameter planes, again
// when Young's moduli change): V lambda, V mu, then V mu (g_a . g_b) for the
// 10 stored pairs. The slot gather reads a contribution's factors from one
// record instead of four parameter planes.
enum StiffRecord
{
    KD_VLAM = 0,
    KD_VMU = 1,
    KD_VMUG = 2,
    KD_COUNT = 12
};

// Inspection record of one tet (pk_system_get_elements / get_rotations):
// the 10 blocks (row-major), forces, K v, R (row-major) and the polar kind.
enum ElemRecord
{
    EREC_K = 0,
    EREC_F = 90,
    EREC_KV = 102,
    EREC_R = 114,
    EREC_KIND = 123,
    EREC_COUNT = 124
};

// Packed gather entry for one upper slot: tet << 5 | pair << 1 | transpose.
inline int pack_contrib(int tet, int pair, int transpose) { return (tet << 5) | (pair << 1) | transpose; }
// Packed rhs entry for one dynamic row: tet << 2 | corner.
inline int pack_corner(int tet, int corner) { return (tet << 2) | corner; }

// Local pair (a, b) -> index of the stored block with min(a,b) <= max(a,b).
extern const int kPairIndex[4][4];

struct Topology
{
    int N = 0, T = 0, G = 1;
    std::vector<int> tets;         // 4T
    std::vector<uint8_t> live;     // T: the tet produces an element (active, valid, V > 0)
    std::vector<uint8_t> tetActive;  // T: the tet's active flag
    std::vector<uint8_t> listed;     // T: live when the gather lists were built (in them)
    std::vector<double> tetShare;    // T: the lumped mass each corner gets from the tet while live
    std::vector<double> prm;       // P_COUNT * T
    std::vector<double> poisson;   // T, kept to refresh Lame parameters when E changes
    std::vector<double> mass;      // N lumped masses (fixed nodes get none)
    std::vector<int> dyn;          // N -> dynamic row or -1
    int nd = 0;
    std::vector<int> rowNode;      // nd
    std::vector<int> groupRowBegin, groupRowEnd;  // G
    int massErrorNode = -1;        // first free active node without mass, or -1

    std::vector<int> ustart, ucol;       // symmetric-upper pattern (nd+1, nnzu)
    std::vector<int> fstart, fcol;       // full pattern (nd+1, nnzf)
    std::vector<int> uFull, uMirror;     // upper slot -> full index of (r,c) and of (c,r) (-1 on the diagonal)
    std::vector<int> cstart, clist;      // upper slot -> packed contributions, increasing tet
    std::vector<int> nstart, nlist;      // dynamic row -> packed (tet, corner), increasing tet
    std::vector<int> cpos;               // 4 T: row-list index of corner a of tet t, or -1
    std::vector<int> astart, alist;      // every node -> packed (tet, corner) over live tets, increasing tet
    std::vector<int> apos;               // 4 T: alist index of corner a of tet t, or -1
    std::vector<int> nodeGroup;          // N -> body
    std::vector<uint8_t> nodeActive;     // N
    std::vector<uint8_t> nodeFixed;      // N
    std::vector<int> scatter;            // optional T*16 upper slot of each local pair, -1 if not dynamic

    long long nnzu() const { return static_cast<long long>(ucol.size()); }
    long long nnzf() const { return static_cast<long long>(fcol.size()); }
};

// Fills every field of topo from desc. Returns PK_OK or PK_ERR_ARG with a
// message in err.
int build_topology(const pk_mesh_desc& desc, Topology& topo, std::string& err, bool wantScatter);

// Rest data and Lame parameters of tets [0, T) into prm (size P_COUNT * T).
void compute_tet_params(const pk_mesh_desc& desc, std::vector<double>& prm);
// Lame parameters from (E, nu) as FEMModel.cpp:174-200 sanitises them.
void lame(double young, double poisson, double* lambda, double* mu);

}  // namespace pk
