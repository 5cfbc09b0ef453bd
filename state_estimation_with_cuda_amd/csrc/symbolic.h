// symbolic.h -- host symbolic analysis of one estimator instance.
//
// Everything here depends only on the topology (closed branches) and on the
// set of active measurements; it is computed once per topology and reused by
// every Gauss-Newton iteration and every solve until the topology changes.
//
// Reference counterparts (all replaced, none translated):
//   Ybus pattern          sparse_matrix.cu:498-626 (atomic-order, duplicates);
//                         here sorted, parallel branches merged
//   H pattern             sparse_matrix.cu:1552-1616 (fixed stride 10); here
//                         exact rows with sorted columns
//   G = H^T W H symbolic  sparse_matrix.cu:1209-1356 (cuSPARSE SpGEMM redone
//                         every call); here once, with a per-entry product list
//   ordering/etree/
//   supernodes            cholesky_optimized.cu:281-387, 809-856, 978-1024
//                         (never instantiated by the reference)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace sle {

// Flat model arrays in internal (insertion) indices.  Branch admittances are
// already computed (network_model.cu:595-612 semantics, FP64).
struct ModelView {
    int n = 0, nb = 0, m = 0, slack = 0;
    const int32_t* bus_type = nullptr;
    const int32_t* from = nullptr;
    const int32_t* to = nullptr;
    const uint8_t* status = nullptr;  // 1 = closed (islands and angle references)
    // Structure: the patterns hold every present branch and measurement, open
    // or closed, so that a breaker operation changes values only (an open
    // branch contributes zero admittance, its measurements zero rows).
    const uint8_t* present = nullptr;    // per branch; nullptr = all present
    const uint8_t* m_present = nullptr;  // per measurement; nullptr = all present
    const int32_t* mtype = nullptr;
    const int32_t* mloc = nullptr;
    const int32_t* mend = nullptr;
    const uint8_t* active = nullptr;
};

enum : int { MT_V_MAG = 0, MT_V_ANGLE = 1, MT_P_INJ = 2, MT_Q_INJ = 3, MT_P_FLOW = 4, MT_Q_FLOW = 5, MT_I_MAG = 6, MT_P_PSEUDO = 7, MT_Q_PSEUDO = 8 };

inline bool is_bus_measurement(int t) {
    return t == MT_V_MAG || t == MT_V_ANGLE || t == MT_P_INJ || t == MT_Q_INJ || t == MT_P_PSEUDO || t == MT_Q_PSEUDO;
}

struct Symbolic {
    int n = 0, nb = 0, m = 0, slack = 0, ns = 0;
    // angle references (one per island of closed branches) and the state layout:
    // th_state[b] = angle state of bus b (-1 for a reference), V_b = nth + b
    std::vector<int32_t> th_state;
    int nth = 0, n_refs = 0;

    // Ybus (CSR, columns sorted, parallel branches merged; closed branches only)
    std::vector<int32_t> yrow, ycol, ydiag;
    std::vector<int32_t> ycptr, ycon;  // per entry: contributions, branch*4 + kind (0 ff, 1 tt, 2 ft, 3 tf)

    // H (CSR over measurement rows, natural state order, sorted); inactive rows empty
    std::vector<int32_t> hrow, hcol;

    // ordering: bus_perm[k] = k-th eliminated bus; perm[k'] = natural state of permuted k'
    std::vector<int32_t> bus_perm, perm, iperm;
    std::vector<int32_t> bus_pth, bus_pv;  // permuted state index of theta_b (-1 for slack) and V_b

    // elimination tree and column counts on permuted states
    std::vector<int32_t> parent, colcount;

    // supernodes (contiguous permuted columns)
    int nsn = 0;
    std::vector<int32_t> sn_first, sn_k, sn_c;  // first column, width, front size
    std::vector<int64_t> sn_rows_off;           // into sn_rows
    std::vector<int32_t> sn_rows;               // permuted row indices of each front (first k = own columns)
    std::vector<int32_t> sn_parent;
    std::vector<int32_t> sn_child_ptr, sn_child;
    std::vector<int64_t> sn_foff;  // dense front (c x c, column-major) offsets; total ftotal
    std::vector<int64_t> sn_moff;  // solve matrix (c x k, column-major) offsets; total mtotal
    std::vector<int64_t> sn_uoff;  // update-vector offsets for the triangular solve; total utotal
    std::vector<int64_t> sn_rmoff; // relative map (child below-rows -> parent front rows), length c-k
    std::vector<int32_t> rmap;
    std::vector<int32_t> sn_level;
    std::vector<int32_t> level_ptr, level_sn;  // supernodes grouped by level (height above leaves)

    // G = H^T W H, lower triangle of the permuted matrix, entries ordered by
    // (supernode, column, row) so the factor reads them contiguously
    int64_t nnz_g = 0;
    std::vector<int32_t> g_row, g_col;  // permuted indices, row >= col
    std::vector<int32_t> g_sn_ptr;      // per supernode range of G entries
    std::vector<int32_t> gloc;          // offset within the front: col_local * c + row_local
    std::vector<int32_t> gpair_ptr;     // per G entry: range of H-position pairs
    std::vector<int32_t> gpair;         // 2 ints per product term: positions in H values
    std::vector<int32_t> gdiag;         // per permuted column: G entry of the diagonal (-1 if none)

    // H^T map in permuted state order (right-hand side H^T W r)
    std::vector<int32_t> tptr, tpos, trow;

    int64_t nnz_l = 0, ftotal = 0, mtotal = 0, utotal = 0, flops = 0;
    int max_c = 0, n_levels = 0;
    double seconds = 0;
};

// Full symbolic analysis.  Throws std::runtime_error on inconsistent input.
void analyze(const ModelView& mv, Symbolic& S);

// Angle state of every bus (-1 for the reference of its island) for the
// current breaker status; equal to Symbolic::th_state while the islands keep
// their references.  islands (optional): per bus, the smallest bus index of
// its island of closed branches.
std::vector<int32_t> angle_states(const ModelView& mv, std::vector<int32_t>* islands = nullptr);

// Natural-order G pattern (CSC, lower), for reporting and parity checks.
void natural_gain_pattern(const Symbolic& S, std::vector<int32_t>& colptr, std::vector<int32_t>& row,
                          std::vector<int64_t>* entry_of = nullptr);

}  // namespace sle
