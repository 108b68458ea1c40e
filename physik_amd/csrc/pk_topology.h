// Host-side topology of the implicit system: everything that depends only on
// connectivity, fixed/active flags and material, built once per change and
// uploaded. Deterministic integer work (the BSR pattern and the element->slot
// gather lists are bit-exact functions of the inputs).
#pragma once

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
    P_COUNT = 22
};

// Element record written by the element kernel, one per tet.
enum ElemLayout
{
    E_BLOCKS = 0,     // 10 blocks R K_ab R^T (a <= b), 9 values each, row-major
    E_FORCE = 90,     // 4 x 3 element forces (elastic + damping)
    E_KV = 102,       // 4 x 3 sum_b (R K_ab R^T) v_b
    E_STRIDE = 114
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
    std::vector<int> astart, alist;      // every node -> packed (tet, corner) over live tets, increasing tet
    std::vector<int> nodeGroup;          // N -> body
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
