// Interpreter for the L2/L3/L4/L6 stencil coefficient tables (TEST
// INFRASTRUCTURE: part of the CPU oracle, never linked into the product).
//
// A table lists, per (term, boundary case), entries (component, coefficient,
// input_a[, input_b]) over the 125 per-site inputs of the reference's
// lcForce::bulkLkForce / boundaryLkForce functions (src/forces/lcForces.h:66-1043):
// q[5], the Q of the 6 neighbours [xd,xu,yd,yu,zd,zu][5] and their first
// derivative vectors [6][15].  The value of one term is L * sum(entries).
#pragma once
#include <vector>

namespace oracle {

struct StencilTerm {
    int comp;
    double coef;
    int a;
    int b;  // -1 for a linear entry
};

// index 0 is the bulk form, 1 + k for boundary case code CASES[k]
struct StencilTables {
    // term slot: 0 -> L2, 1 -> L3, 2 -> L4, 3 -> L6
    std::vector<StencilTerm> t[4][28];
    bool loaded = false;
};

// boundary case code xx + 4*yy + 16*zz -> table slot (lcForces.h:13-25)
inline int caseSlot(int code) {
    int xx = code & 3, yy = (code >> 2) & 3, zz = (code >> 4) & 3;
    return 1 + xx + 3 * yy + 9 * zz;
}

inline void evalStencil(const std::vector<StencilTerm>& terms, double L, const double* x, double* out) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (const StencilTerm& e : terms) {
        if (e.b < 0)
            acc[e.comp] += e.coef * x[e.a];
        else
            acc[e.comp] += e.coef * x[e.a] * x[e.b];
    }
    for (int r = 0; r < 5; ++r) out[r] = L * acc[r];
}

}  // namespace oracle
