/*
 * ORACLE — test infrastructure, not product code.
 *
 * Floating-point operations per cell of one substep of the reference's MHD solver, counted by
 * running the restated right-hand side (oracle.h: rhs_cell, modelsolver.c:880-912) and the
 * two-buffer RK3 update of the GPU path (oracle.h: gpu_form_substep, integration.cuh:14-37) on a
 * scalar type that records its own arithmetic as a graph. Operations whose operands are all
 * constants (the stencil coefficients, folded at compile time) or that add a constant zero (the
 * first term of an accumulation) are not recorded. Two counts are reported: every recorded
 * operation ("all": what the host formulation executes, which evaluates all six second
 * derivatives of every field), and the operations the eight rates depend on ("live": what is
 * left once unused results are dropped, as a compiler drops them).
 *
 * Prints "<what> <all|live> add mul div exp neg" lines: "rhs" for the right-hand side of all
 * eight fields, "update0" / "update12" for the RK3 update of the eight fields on substep 0 and
 * on substeps 1-2, and "first", "second", "cross" for one derivative of each kind (checks of the
 * counter against their closed forms).
 */
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace {

enum Op { ADD, MUL, DIV, EXP, NEG, NOPS };
struct GraphNode {
    Op op;
    int a, b; // operand nodes (-1: an input or a constant)
};
std::vector<GraphNode> g_nodes;

} // namespace

struct Counted {
    double v;
    bool k; // a constant: known when the code is compiled
    int id; // graph node that produced the value (-1: an input or a constant)
    Counted() : v(0), k(true), id(-1) {}
    Counted(double x) : v(x), k(true), id(-1) {}
    Counted(double x, bool konst) : v(x), k(konst), id(-1) {}
};

static inline Counted
record(Op op, double v, const Counted& a, const Counted& b, bool counted)
{
    Counted r(v, a.k && b.k);
    if (counted) {
        r.id = (int)g_nodes.size();
        g_nodes.push_back({op, a.id, b.id});
    }
    else if (!r.k) { // x + 0, 0 + x: the value of the non-zero operand
        r.id = (a.k && a.v == 0) ? b.id : a.id;
    }
    return r;
}
static inline bool
free_op(const Counted& a, const Counted& b)
{
    return (a.k && b.k) || (a.k && a.v == 0) || (b.k && b.v == 0);
}
static inline Counted
operator+(Counted a, Counted b)
{
    return record(ADD, a.v + b.v, a, b, !free_op(a, b));
}
static inline Counted
operator-(Counted a, Counted b)
{
    return record(ADD, a.v - b.v, a, b, !free_op(a, b));
}
static inline Counted
operator*(Counted a, Counted b)
{
    return record(MUL, a.v * b.v, a, b, !(a.k && b.k));
}
static inline Counted
operator/(Counted a, Counted b)
{
    return record(DIV, a.v / b.v, a, b, !(a.k && b.k));
}
static inline Counted
operator-(Counted a)
{
    return record(NEG, -a.v, a, Counted(), !a.k);
}
static inline Counted& operator+=(Counted& a, Counted b) { return a = a + b; }
static inline Counted& operator-=(Counted& a, Counted b) { return a = a - b; }
static inline Counted& operator*=(Counted& a, Counted b) { return a = a * b; }
static inline Counted& operator/=(Counted& a, Counted b) { return a = a / b; }
static inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
static inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
static inline Counted
exp(Counted a)
{
    return record(EXP, std::exp(a.v), a, Counted(), !a.k);
}

// Counts of the nodes recorded since `first`: all of them, and those that `outs` depend on.
static void
report(const char* what, size_t first, const Counted* outs, int nouts)
{
    long all[NOPS] = {}, live[NOPS] = {};
    std::vector<char> mark(g_nodes.size(), 0);
    std::vector<int> todo;
    for (int i = 0; i < nouts; ++i)
        if (outs[i].id >= 0)
            todo.push_back(outs[i].id);
    while (!todo.empty()) {
        const int n = todo.back();
        todo.pop_back();
        if (n < 0 || mark[n])
            continue;
        mark[n] = 1;
        todo.push_back(g_nodes[n].a);
        todo.push_back(g_nodes[n].b);
    }
    for (size_t n = first; n < g_nodes.size(); ++n) {
        ++all[g_nodes[n].op];
        if (mark[n])
            ++live[g_nodes[n].op];
    }
    printf("%s all %ld %ld %ld %ld %ld\n", what, all[ADD], all[MUL], all[DIV], all[EXP], all[NEG]);
    printf("%s live %ld %ld %ld %ld %ld\n", what, live[ADD], live[MUL], live[DIV], live[EXP],
           live[NEG]);
}

int
main()
{
    using namespace oracle;
    const int n = 8;
    const Grid<Counted> g(n, n, n);
    double rp[P_NUM] = {};
    rp[P_dsx] = rp[P_dsy] = rp[P_dsz] = 0.1;
    rp[P_nu_visc] = rp[P_eta] = 5e-3;
    rp[P_cs_sound] = 1.0;
    rp[P_mu0] = 1.4, rp[P_cp_sound] = 1.0, rp[P_gamma] = 5.0 / 3.0;
    rp[P_lnT0] = 1.2, rp[P_lnrho0] = 1.3, rp[P_zeta] = 0.01;
    Params<Counted> p = make_params<Counted>(rp);
    // the parameters are uniforms: values at run time, not constants
    for (Counted* q : {&p.inv_dsx, &p.inv_dsy, &p.inv_dsz, &p.nu_visc, &p.eta, &p.mu0, &p.cp_sound,
                       &p.gamma, &p.lnT0, &p.lnrho0, &p.zeta, &p.cs2_sound})
        q->k = false;

    std::vector<Counted> in((size_t)NFIELDS * g.size());
    Counted* fin[NFIELDS];
    for (int q = 0; q < NFIELDS; ++q) {
        fin[q] = in.data() + (size_t)q * g.size();
        for (size_t i = 0; i < g.size(); ++i)
            fin[q][i] = Counted(1e-3 * (double)((i * 7 + q * 13) % 17), false);
    }

    const int c = NGHOST_ + n / 2;
    Counted rate[NFIELDS];
    size_t first = g_nodes.size();
    rhs_cell(g, fin, c, c, c, p, rate);
    report("rhs", first, rate, NFIELDS);

    // one cell's update: the substep over one cell with its right-hand side made an input
    const Counted dt(1e-3, false);
    const Counted alpha[] = {0.0, -5. / 9., -153. / 128.}, beta[] = {1. / 3., 15. / 16., 8. / 15.};
    for (int step = 0; step < 2; ++step) {
        Counted res[NFIELDS];
        first = g_nodes.size();
        for (int q = 0; q < NFIELDS; ++q) {
            const Counted r(rate[q].v, false), cur(1e-3 * (q + 1), false), prev(2e-3 * (q + 1), false);
            // the expression of gpu_form_substep (integration.cuh:14-37)
            res[q] = step == 0 ? cur + beta[0] * r * dt
                               : cur + beta[1] * (alpha[1] * (Counted(1.) / beta[0]) * (cur - prev) +
                                                  r * dt);
        }
        report(step == 0 ? "update0" : "update12", first, res, NFIELDS);
    }

    Counted pa[7], pb[7];
    for (int o = 0; o < 7; ++o) {
        pa[o] = Counted(0.1 * o, false);
        pb[o] = Counted(0.2 * o, false);
    }
    Counted d;
    first = g_nodes.size();
    d     = first_derivative(pa, p.inv_dsx);
    report("first", first, &d, 1);
    first = g_nodes.size();
    d     = second_derivative(pa, p.inv_dsx);
    report("second", first, &d, 1);
    first = g_nodes.size();
    d     = cross_derivative(pa, pb, p.inv_dsx, p.inv_dsy);
    report("cross", first, &d, 1);
    return 0;
}
