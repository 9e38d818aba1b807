// Host-side exercise of the library and the oracle under AddressSanitizer
// and UndefinedBehaviorSanitizer (built by tests/test_sanitizers_cpu.py with
// the sanitizers on host code only; no GPU is needed or touched).
//
// Library: every C-ABI function that works without a device -- the fused-pass
// planner on the 28-qubit benchmark circuit and on random circuits, the
// host gate compiler for every gate, the Pauli-word planner, batched
// templates (compiled on the host, uploaded only when run), the error paths
// of argument checking.  Oracle: gates, generators, expectation values,
// CSR, probabilities, the sampler and the adjoint Jacobian on small states.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "lgh.h"

// The oracle's C interface (oracle/oracle.h, C99 complex); complex arrays
// cross as untyped pointers here.
extern "C" {
int orc_apply_op(int n, void *psi, const char *name, const int *wires, int nw, int inverse, const double *params,
                 const void *matrix);
int orc_apply_ops(int n, void *psi, int nops, const char *const *names, const int *wires_flat, const int *woff,
                  const int *inverses, const double *params, const void *const *matrices);
int orc_probs(int n, const void *psi, const int *wires, int k, double *out);
int orc_sample(int n, const void *psi, uint64_t shots, uint64_t seed, uint64_t *out);
int orc_expval_pauli_words(int n, const void *psi, int nwords, const char *paulis, const int *wires_flat,
                           const int64_t *offsets, const void *coeffs, double *per_word, double *total);
int orc_adjoint_jacobian(int n, const void *psi, int nobs, const void *const *bras, int nops,
                         const char *const *names, const int *wires_flat, const int *woff, const int *inverses,
                         const double *params, const int *nparams, const void *const *matrices,
                         const int64_t *trainable, int ntp, double *jac);
}

static int failures = 0;
#define EXPECT(c)                                                   \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "EXPECT failed line %d: %s\n", __LINE__, #c); \
            ++failures;                                             \
        }                                                           \
    } while (0)

struct Circuit {
    std::vector<lgh_op> ops;
    std::vector<std::vector<int32_t>> wires;
    void add(int gate, std::vector<int32_t> w, double p = 0.0, bool param = false, bool inv = false) {
        wires.push_back(std::move(w));
        lgh_op o{};
        o.gate = gate;
        o.num_wires = int32_t(wires.back().size());
        o.inverse = inv;
        o.params[0] = p;
        o.num_params = param ? 1 : 0;
        ops.push_back(o);
    }
    void fix() {
        for (size_t i = 0; i < ops.size(); ++i) ops[i].wires = wires[i].data();
    }
};

static Circuit brick(int n, int layers, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0, 6.283);
    Circuit c;
    for (int l = 0; l < layers; ++l) {
        for (int w = 0; w < n; ++w) {
            c.add(LGH_GATE_RY, {w}, u(rng), true);
            c.add(LGH_GATE_RZ, {w}, u(rng), true, w % 3 == 0);
        }
        for (int s = 0; s < 2; ++s)
            for (int w = s; w + 1 < n; w += 2) c.add(LGH_GATE_CZ, {w, w + 1});
    }
    c.fix();
    return c;
}

static Circuit random_circuit(int n, int len, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> g(0, LGH_GATE_MULTIRZ), wd(0, n - 1);
    std::uniform_real_distribution<double> u(-3, 3);
    Circuit c;
    for (int i = 0; i < len; ++i) {
        int gate = g(rng);
        // multi-parameter gates have no adjoint form (the sweep rejects them)
        if (gate == LGH_GATE_ROT || gate == LGH_GATE_CROT) gate = LGH_GATE_RY;
        int nw = 1;
        switch (gate) {
        case LGH_GATE_CNOT: case LGH_GATE_SWAP: case LGH_GATE_CY: case LGH_GATE_CZ: case LGH_GATE_CONTROLLEDPHASESHIFT:
        case LGH_GATE_CRX: case LGH_GATE_CRY: case LGH_GATE_CRZ: case LGH_GATE_CROT: case LGH_GATE_ISINGXX:
        case LGH_GATE_ISINGYY: case LGH_GATE_ISINGZZ: case LGH_GATE_SINGLEEXCITATION:
        case LGH_GATE_SINGLEEXCITATIONMINUS: case LGH_GATE_SINGLEEXCITATIONPLUS: nw = 2; break;
        case LGH_GATE_CSWAP: case LGH_GATE_TOFFOLI: nw = 3; break;
        case LGH_GATE_DOUBLEEXCITATION: case LGH_GATE_DOUBLEEXCITATIONMINUS:
        case LGH_GATE_DOUBLEEXCITATIONPLUS: nw = 4; break;
        case LGH_GATE_MULTIRZ: nw = 1 + i % 4; break;
        default: nw = 1;
        }
        std::vector<int32_t> w;
        while (int(w.size()) < nw) {
            const int x = wd(rng);
            bool dup = false;
            for (int y : w) dup = dup || y == x;
            if (!dup) w.push_back(x);
        }
        const bool param = gate == LGH_GATE_RX || gate == LGH_GATE_RY || gate == LGH_GATE_RZ ||
                           gate == LGH_GATE_PHASESHIFT || gate >= LGH_GATE_CRX;
        c.add(gate, w, u(rng), param, i % 5 == 0);
    }
    c.fix();
    return c;
}

static void library_checks() {
    // fused-pass planner
    Circuit b = brick(28, 6, 1337);
    int32_t passes = 0, standalone = 0, runs = 0, members = 0;
    EXPECT(lgh_fusion_plan(28, b.ops.data(), int64_t(b.ops.size()), 0, LGH_C128, &passes, &standalone, &runs,
                           &members) == LGH_OK);
    EXPECT(passes > 0 && passes < 20);
    EXPECT(lgh_fusion_plan(28, b.ops.data(), int64_t(b.ops.size()), 1, LGH_C128, &passes, &standalone, &runs,
                           &members) == LGH_OK);
    for (unsigned seed = 0; seed < 12; ++seed) {
        const int n = 9 + int(seed % 9);
        Circuit r = random_circuit(n, 60, seed);
        for (int adj = 0; adj < 2; ++adj)
            for (lgh_dtype dt : {LGH_C64, LGH_C128})
                EXPECT(lgh_fusion_plan(n, r.ops.data(), int64_t(r.ops.size()), adj, dt, &passes, &standalone, &runs,
                                       &members) == LGH_OK);
        // host gate compiler
        for (const lgh_op &o : r.ops) {
            std::vector<double> m(size_t(2) << (2 * o.num_wires));
            EXPECT(lgh_gate_matrix(&o, m.data()) == LGH_OK);
        }
        Circuit rot;
        rot.add(LGH_GATE_ROT, {n - 1}, 0.3, true);
        rot.add(LGH_GATE_CROT, {0, n - 1}, -0.7, true, true);
        rot.fix();
        for (lgh_op &o : rot.ops) {
            o.params[1] = 1.1;
            o.params[2] = -0.2;
            o.num_params = 3;
            std::vector<double> m(size_t(2) << (2 * o.num_wires));
            EXPECT(lgh_gate_matrix(&o, m.data()) == LGH_OK);
        }
        EXPECT(lgh_fusion_plan(n, rot.ops.data(), 2, 0, LGH_C128, &passes, &standalone, &runs, &members) == LGH_OK);
        // the adjoint rejects multi-parameter gates, as the reference does
        EXPECT(lgh_fusion_plan(n, rot.ops.data(), 2, 1, LGH_C128, &passes, &standalone, &runs, &members) ==
               LGH_ERR_UNSUPPORTED || std::strlen(lgh_last_error()) > 0);
    }
    // Pauli-word planner: words inside tiles, words wider than a tile
    {
        std::mt19937 rng(7);
        std::string paulis;
        std::vector<int32_t> wires;
        std::vector<int64_t> offs{0};
        for (int w = 0; w < 300; ++w) {
            const int k = w % 50 == 0 ? 20 : 1 + int(rng() % 4);
            std::vector<int32_t> ws;
            while (int(ws.size()) < k) {
                const int x = int(rng() % 24);
                bool dup = false;
                for (int y : ws) dup = dup || y == x;
                if (!dup) ws.push_back(x);
            }
            for (int x : ws) {
                paulis.push_back(k == 20 ? 'X' : "IXYZ"[rng() % 4]);
                wires.push_back(x);
            }
            offs.push_back(int64_t(paulis.size()));
        }
        int32_t tp = 0, tl = 0, pl = 0;
        EXPECT(lgh_pauli_plan(24, 300, paulis.data(), wires.data(), offs.data(), &tp, &tl, &pl) == LGH_OK);
        EXPECT(tl >= tp && tp > 0 && pl > 0);
        EXPECT(lgh_pauli_plan(3, 300, paulis.data(), wires.data(), offs.data(), &tp, &tl, &pl) ==
               LGH_ERR_INVALID_ARGUMENT); // wires out of range for 3 qubits
    }
    // batched templates: compiled on the host, uploaded only when run
    for (int n : {3, 9, 14, 20}) {
        Circuit r = brick(n, 3, unsigned(n));
        std::vector<int32_t> pidx(r.ops.size(), -1);
        int k = 0;
        for (size_t i = 0; i < r.ops.size(); ++i)
            if (r.ops[i].num_params) pidx[i] = k++;
        const int32_t obs_p[2] = {LGH_GATE_PAULIZ, LGH_GATE_PAULIX}, obs_w[2] = {0, n - 1};
        lgh_tmpl *t = nullptr;
        EXPECT(lgh_tmpl_create(&t, n, r.ops.data(), int32_t(r.ops.size()), pidx.data(), obs_p, obs_w, 2, 0) == LGH_OK);
        int32_t np = 0, nq = 0, f = 0, rv = 0, sq = 0;
        EXPECT(lgh_tmpl_info(t, &np, &nq) == LGH_OK && np == k && nq == n);
        EXPECT(lgh_tmpl_passes(t, &f, &rv, &sq) == LGH_OK && f > 0 && rv > 0 && sq >= 9);
        EXPECT(lgh_tmpl_destroy(t) == LGH_OK);
    }
    // argument checks
    Circuit bad;
    bad.add(LGH_GATE_CNOT, {2, 2});
    bad.fix();
    EXPECT(lgh_fusion_plan(4, bad.ops.data(), 1, 0, LGH_C128, &passes, &standalone, &runs, &members) != LGH_OK);
    EXPECT(std::strlen(lgh_last_error()) > 0);
    lgh_op unknown{};
    unknown.gate = 999;
    std::vector<double> m(8);
    EXPECT(lgh_gate_matrix(&unknown, m.data()) != LGH_OK);
    EXPECT(lgh_tmpl_create(nullptr, 4, nullptr, 0, nullptr, nullptr, nullptr, 0, 0) != LGH_OK);
    EXPECT(lgh_sv_create(nullptr, 4, LGH_C128, 0, nullptr) != LGH_OK);
}

static void oracle_checks() {
    typedef std::complex<double> C;
    const int n = 8;
    std::vector<C> psi(size_t(1) << n);
    psi[0] = 1.0;
    Circuit r = random_circuit(n, 40, 99);
    std::vector<const char *> names;
    const char *nm[] = {"Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP", "CSWAP",
                        "Toffoli", "CY", "CZ", "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX",
                        "CRY", "CRZ", "CRot", "IsingXX", "IsingYY", "IsingZZ", "SingleExcitation",
                        "SingleExcitationMinus", "SingleExcitationPlus", "DoubleExcitation",
                        "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ"};
    std::vector<int> wflat, woff{0}, inv, npar;
    std::vector<double> params;
    for (const lgh_op &o : r.ops) {
        names.push_back(nm[o.gate]);
        for (int k = 0; k < o.num_wires; ++k) wflat.push_back(o.wires[k]);
        woff.push_back(int(wflat.size()));
        inv.push_back(o.inverse);
        npar.push_back(o.num_params);
        params.insert(params.end(), o.params, o.params + 3);
    }
    std::vector<const void *> mats(r.ops.size(), nullptr);
    EXPECT(orc_apply_ops(n, static_cast<void *>(psi.data()), int(r.ops.size()), names.data(),
                         wflat.data(), woff.data(), inv.data(), params.data(), mats.data()) == 0);
    double nrm = 0;
    for (const C &z : psi) nrm += std::norm(z);
    EXPECT(std::fabs(nrm - 1.0) < 1e-12);
    // adjoint of <Z_0>
    std::vector<C> bra(psi);
    const int w0 = 0;
    EXPECT(orc_apply_op(n, static_cast<void *>(bra.data()), "PauliZ", &w0, 1, 0, nullptr, nullptr) == 0);
    int ntp = 0;
    for (int p : npar) ntp += p > 0;
    std::vector<int64_t> tp(static_cast<size_t>(ntp));
    for (int k = 0; k < ntp; ++k) tp[size_t(k)] = k;
    std::vector<double> jac(static_cast<size_t>(ntp));
    const void *bras[1] = {static_cast<const void *>(bra.data())};
    EXPECT(orc_adjoint_jacobian(n, static_cast<const void *>(psi.data()), 1, bras, int(r.ops.size()),
                                names.data(), wflat.data(), woff.data(), inv.data(), params.data(), npar.data(),
                                mats.data(), tp.data(), ntp, jac.data()) == 0);
    // measurements and the sampler
    const int pw[3] = {0, 5, 7};
    std::vector<double> pr(8);
    EXPECT(orc_probs(n, static_cast<const void *>(psi.data()), pw, 3, pr.data()) == 0);
    std::vector<uint64_t> smp(5000);
    EXPECT(orc_sample(n, static_cast<const void *>(psi.data()), smp.size(), 1337, smp.data()) == 0);
    for (size_t i = 1; i < smp.size(); ++i) EXPECT(smp[i - 1] <= smp[i]);
    const char *pw_chars = "XYZZ";
    const int pw_wires[4] = {1, 2, 3, 6};
    const int64_t pw_offs[3] = {0, 2, 4};
    const C coefs[2] = {0.5, -1.0};
    double per[2], tot = 0;
    EXPECT(orc_expval_pauli_words(n, static_cast<const void *>(psi.data()), 2, pw_chars, pw_wires,
                                  pw_offs, static_cast<const void *>(coefs), per, &tot) == 0);
}

int main() {
    library_checks();
    oracle_checks();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host checks passed\n");
    return 0;
}
