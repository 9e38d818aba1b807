// StateVectorHIPManaged: host class of the MI355X statevector, the
// counterpart of StateVectorCudaManaged
// (pennylane_lightning_gpu/src/simulator/StateVectorCudaManaged.hpp:69-1714
// and StateVectorCudaBase.hpp:54-313).  It owns one device-resident state
// through the C ABI in include/lgh.h and exposes the method set the pybind11
// module binds (Bindings.cpp:82-610).  No HIP header is included here: all
// device work happens behind the C ABI.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "lgh.h"

namespace Pennylane::HIP {

// Error type surfaced to Python as PLException (Bindings.cpp:1734).
class LightningException : public std::exception {
  public:
    explicit LightningException(std::string msg) : msg_(std::move(msg)) {}
    const char *what() const noexcept override { return msg_.c_str(); }

  private:
    std::string msg_;
};

inline void check(lgh_status s) {
    if (s != LGH_OK) throw LightningException(lgh_last_error());
}

#define PL_HIP_ABORT_IF(cond, msg)                                                                      \
    do {                                                                                               \
        if (cond) throw ::Pennylane::HIP::LightningException(msg);                                     \
    } while (0)

// Device + stream tag (DevTag.hpp:14-56).  A null stream selects the
// library's own stream for the device.
class DevTag {
  public:
    DevTag() = default;
    explicit DevTag(int device_id, void *stream = nullptr) : device_(device_id), stream_(stream) {}
    int getDeviceID() const { return device_; }
    void *getStreamID() const { return stream_; }
    void refresh() {}

  private:
    int device_ = 0;
    void *stream_ = nullptr;
};

inline std::size_t getGPUCount() {
    int c = 0;
    check(lgh_device_count(&c));
    return static_cast<std::size_t>(c);
}

// Gate-name table: the reference's named gate methods (Bindings.cpp:120-427).
inline const std::unordered_map<std::string, int> &gateIds() {
    static const std::unordered_map<std::string, int> ids = {
        {"Identity", LGH_GATE_IDENTITY},
        {"PauliX", LGH_GATE_PAULIX},
        {"PauliY", LGH_GATE_PAULIY},
        {"PauliZ", LGH_GATE_PAULIZ},
        {"Hadamard", LGH_GATE_HADAMARD},
        {"S", LGH_GATE_S},
        {"T", LGH_GATE_T},
        {"CNOT", LGH_GATE_CNOT},
        {"SWAP", LGH_GATE_SWAP},
        {"CSWAP", LGH_GATE_CSWAP},
        {"Toffoli", LGH_GATE_TOFFOLI},
        {"CY", LGH_GATE_CY},
        {"CZ", LGH_GATE_CZ},
        {"PhaseShift", LGH_GATE_PHASESHIFT},
        {"ControlledPhaseShift", LGH_GATE_CONTROLLEDPHASESHIFT},
        {"RX", LGH_GATE_RX},
        {"RY", LGH_GATE_RY},
        {"RZ", LGH_GATE_RZ},
        {"Rot", LGH_GATE_ROT},
        {"CRX", LGH_GATE_CRX},
        {"CRY", LGH_GATE_CRY},
        {"CRZ", LGH_GATE_CRZ},
        {"CRot", LGH_GATE_CROT},
        {"IsingXX", LGH_GATE_ISINGXX},
        {"IsingYY", LGH_GATE_ISINGYY},
        {"IsingZZ", LGH_GATE_ISINGZZ},
        {"SingleExcitation", LGH_GATE_SINGLEEXCITATION},
        {"SingleExcitationMinus", LGH_GATE_SINGLEEXCITATIONMINUS},
        {"SingleExcitationPlus", LGH_GATE_SINGLEEXCITATIONPLUS},
        {"DoubleExcitation", LGH_GATE_DOUBLEEXCITATION},
        {"DoubleExcitationMinus", LGH_GATE_DOUBLEEXCITATIONMINUS},
        {"DoubleExcitationPlus", LGH_GATE_DOUBLEEXCITATIONPLUS},
        {"MultiRZ", LGH_GATE_MULTIRZ},
    };
    return ids;
}

inline int gateId(const std::string &name) {
    auto it = gateIds().find(name);
    return it == gateIds().end() ? -1 : it->second;
}

inline bool isStatePrep(const std::string &name) {
    return name == "QubitStateVector" || name == "StatePrep" || name == "BasisState";
}

// Builds lgh_op records; owns the wire and matrix storage they point into.
class OpBuilder {
  public:
    template <class PrecisionT>
    void add(const std::string &name, const std::vector<std::size_t> &wires, bool inverse,
             const std::vector<PrecisionT> &params, const std::vector<std::complex<PrecisionT>> &matrix) {
        lgh_op op{};
        const int id = gateId(name);
        wires_.emplace_back(wires.begin(), wires.end());
        if (isStatePrep(name)) {
            op.gate = LGH_GATE_STATEPREP;
        } else if (id >= 0) {
            op.gate = id;
        } else if (!matrix.empty()) {
            op.gate = LGH_GATE_MATRIX;
            std::vector<double> m(2 * matrix.size());
            for (std::size_t i = 0; i < matrix.size(); ++i) {
                m[2 * i] = static_cast<double>(matrix[i].real());
                m[2 * i + 1] = static_cast<double>(matrix[i].imag());
            }
            std::size_t dim = 1;
            while (dim * dim < matrix.size()) dim <<= 1;
            PL_HIP_ABORT_IF(dim * dim != matrix.size() || dim != (std::size_t(1) << wires.size()),
                            "Matrix size does not match the number of wires of " + name);
            mats_.push_back(std::move(m));
            mat_index_.push_back(mats_.size() - 1);
        } else {
            throw LightningException("Currently unsupported gate: " + name);
        }
        if (op.gate != LGH_GATE_MATRIX) mat_index_.push_back(SIZE_MAX);
        op.num_wires = static_cast<int32_t>(wires.size());
        op.inverse = inverse ? 1 : 0;
        for (std::size_t i = 0; i < params.size() && i < 3; ++i) op.params[i] = static_cast<double>(params[i]);
        op.num_params = static_cast<int32_t>(params.size());
        const int needs = (op.gate == LGH_GATE_ROT || op.gate == LGH_GATE_CROT) ? 3
                          : (op.gate == LGH_GATE_RX || op.gate == LGH_GATE_RY || op.gate == LGH_GATE_RZ ||
                             op.gate == LGH_GATE_PHASESHIFT || op.gate == LGH_GATE_CONTROLLEDPHASESHIFT ||
                             op.gate == LGH_GATE_CRX || op.gate == LGH_GATE_CRY || op.gate == LGH_GATE_CRZ ||
                             op.gate == LGH_GATE_ISINGXX || op.gate == LGH_GATE_ISINGYY ||
                             op.gate == LGH_GATE_ISINGZZ || op.gate == LGH_GATE_MULTIRZ ||
                             (op.gate >= LGH_GATE_SINGLEEXCITATION && op.gate <= LGH_GATE_DOUBLEEXCITATIONPLUS))
                              ? 1
                              : 0;
        PL_HIP_ABORT_IF(static_cast<int>(params.size()) < needs, "Missing parameters for " + name);
        ops_.push_back(op);
    }

    const lgh_op *data() {
        // wire/matrix storage is stable now: bind the pointers
        for (std::size_t i = 0; i < ops_.size(); ++i) {
            ops_[i].wires = wires_[i].data();
            ops_[i].matrix = mat_index_[i] == SIZE_MAX ? nullptr : mats_[mat_index_[i]].data();
        }
        return ops_.data();
    }
    std::size_t size() const { return ops_.size(); }
    void clear() {
        ops_.clear();
        wires_.clear();
        mats_.clear();
        mat_index_.clear();
    }

  private:
    std::vector<lgh_op> ops_;
    std::vector<std::vector<int32_t>> wires_;
    std::vector<std::vector<double>> mats_;
    std::vector<std::size_t> mat_index_;
};

template <class PrecisionT> class StateVectorHIPManaged {
  public:
    static_assert(std::is_same_v<PrecisionT, float> || std::is_same_v<PrecisionT, double>);
    using Precision = PrecisionT;
    using ComplexT = std::complex<PrecisionT>;
    static constexpr lgh_dtype kDtype = std::is_same_v<PrecisionT, double> ? LGH_C128 : LGH_C64;

    explicit StateVectorHIPManaged(std::size_t num_qubits, const DevTag &tag = DevTag{}) : tag_(tag) {
        check(lgh_sv_create(&sv_, static_cast<int>(num_qubits), kDtype, tag.getDeviceID(), tag.getStreamID()));
        num_qubits_ = num_qubits;
        initSV();
    }
    // A state whose amplitudes the caller overwrites next (no |0...0> fill).
    struct Uninitialized {};
    StateVectorHIPManaged(std::size_t num_qubits, const DevTag &tag, Uninitialized) : tag_(tag) {
        check(lgh_sv_create(&sv_, static_cast<int>(num_qubits), kDtype, tag.getDeviceID(), tag.getStreamID()));
        num_qubits_ = num_qubits;
    }
    StateVectorHIPManaged(const ComplexT *data, std::size_t length, const DevTag &tag = DevTag{}) : tag_(tag) {
        PL_HIP_ABORT_IF(length == 0 || (length & (length - 1)) != 0, "State length must be a power of two");
        std::size_t n = 0;
        while ((std::size_t(1) << n) < length) ++n;
        check(lgh_sv_create(&sv_, static_cast<int>(n), kDtype, tag.getDeviceID(), tag.getStreamID()));
        num_qubits_ = n;
        CopyHostDataToGpu(data, length);
    }
    StateVectorHIPManaged(const StateVectorHIPManaged &other) : tag_(other.tag_), num_qubits_(other.num_qubits_) {
        check(lgh_sv_create(&sv_, static_cast<int>(num_qubits_), kDtype, tag_.getDeviceID(), tag_.getStreamID()));
        check(lgh_sv_copy_d2d(sv_, other.sv_));
    }
    StateVectorHIPManaged &operator=(const StateVectorHIPManaged &) = delete;
    StateVectorHIPManaged(StateVectorHIPManaged &&o) noexcept
        : sv_(o.sv_), tag_(o.tag_), num_qubits_(o.num_qubits_), seed_(o.seed_), sample_calls_(o.sample_calls_) {
        o.sv_ = nullptr;
    }
    ~StateVectorHIPManaged() {
        if (sv_) (void)lgh_sv_destroy(sv_);
    }

    lgh_sv *handle() const { return sv_; }
    std::size_t getNumQubits() const { return num_qubits_; }
    std::size_t getLength() const { return std::size_t(1) << num_qubits_; }
    const DevTag &getDevTag() const { return tag_; }
    int getDeviceID() const { return tag_.getDeviceID(); }

    // ---- initialisation (StateVectorCudaBase.hpp:234-240, SVCM:144-185)
    void initSV(bool async = false) {
        (void)async;
        check(lgh_sv_reset(sv_));
    }
    void setBasisState(const ComplexT &value, std::size_t index, bool async = false) {
        (void)async;
        check(lgh_sv_set_basis(sv_, index, value.real(), value.imag()));
    }
    template <class IndexT>
    void setStateVector(IndexT num_indices, const ComplexT *values, const IndexT *indices, bool async = false) {
        (void)async;
        std::vector<int64_t> idx(indices, indices + num_indices);
        check(lgh_sv_set_indexed(sv_, idx.data(), values, static_cast<uint64_t>(num_indices)));
    }

    // ---- copies (StateVectorCudaBase.hpp:104-219)
    void CopyHostDataToGpu(const ComplexT *host, std::size_t length, bool async = false) {
        (void)async;
        check(lgh_sv_copy_h2d(sv_, host, length));
    }
    void CopyHostDataToGpu(const std::vector<ComplexT> &host, bool async = false) {
        CopyHostDataToGpu(host.data(), host.size(), async);
    }
    void CopyGpuDataToHost(ComplexT *host, std::size_t length, bool async = false) const {
        (void)async;
        check(lgh_sv_copy_d2h(sv_, host, length));
    }
    std::vector<ComplexT> getDataVector() const {
        std::vector<ComplexT> v(getLength());
        CopyGpuDataToHost(v.data(), v.size());
        return v;
    }
    void updateData(const StateVectorHIPManaged &other, bool async = false) {
        check(lgh_sv_copy_d2d(sv_, other.sv_));
        if (!async) check(lgh_sv_synchronize(sv_));
    }
    void synchronize() const { check(lgh_sv_synchronize(sv_)); }
    // this <- P other for a Pauli word (letters I/X/Y/Z on wires), one pass
    void setPauliWordOf(const StateVectorHIPManaged &other, const std::string &letters,
                        const std::vector<std::size_t> &wires) {
        std::vector<int32_t> w(wires.begin(), wires.end());
        check(lgh_sv_pauli_word_from(sv_, other.sv_, letters.data(), w.data(), static_cast<int32_t>(w.size())));
    }

    // ---- gates (SVCM:198-315)
    void applyOperation(const std::string &opName, const std::vector<std::size_t> &wires, bool adjoint = false,
                        const std::vector<PrecisionT> &params = {0.0}, const std::vector<ComplexT> &matrix = {}) {
        if (opName == "Identity") return;
        OpBuilder b;
        b.add<PrecisionT>(opName, wires, adjoint, params, matrix);
        PL_HIP_ABORT_IF(b.data()[0].gate == LGH_GATE_STATEPREP, "State preparation is not a gate: " + opName);
        check(lgh_apply(sv_, b.data(), 1));
    }
    void applyOperation_std(const std::string &opName, const std::vector<std::size_t> &wires, bool adjoint,
                            const std::vector<PrecisionT> &params, const std::vector<ComplexT> &matrix) {
        applyOperation(opName, wires, adjoint, params, matrix);
    }
    void applyOperation(const std::vector<std::string> &opNames, const std::vector<std::vector<std::size_t>> &wires,
                        const std::vector<bool> &adjoints, const std::vector<std::vector<PrecisionT>> &params) {
        PL_HIP_ABORT_IF(opNames.size() != wires.size(), "Incompatible number of ops and wires");
        PL_HIP_ABORT_IF(opNames.size() != adjoints.size(), "Incompatible number of ops and adjoints");
        PL_HIP_ABORT_IF(opNames.size() != params.size(), "Incompatible number of ops and parameters");
        OpBuilder b;
        for (std::size_t i = 0; i < opNames.size(); ++i) {
            if (opNames[i] == "Identity") continue;
            b.add<PrecisionT>(opNames[i], wires[i], adjoints[i], params[i], {});
        }
        if (b.size()) check(lgh_apply(sv_, b.data(), static_cast<int64_t>(b.size())));
    }
    void applyOperation(const std::vector<std::string> &opNames, const std::vector<std::vector<std::size_t>> &wires,
                        const std::vector<bool> &adjoints) {
        applyOperation(opNames, wires, adjoints, std::vector<std::vector<PrecisionT>>(opNames.size()));
    }
    // Generator of a differentiable gate (GateGenerators.hpp:58-320); returns
    // the scaling factor of the generator (AdjointDiffGPU.hpp:96-114).
    PrecisionT applyGenerator(const std::string &opName, const std::vector<std::size_t> &wires, bool adjoint = false) {
        (void)adjoint; // generators are Hermitian
        OpBuilder b;
        b.add<PrecisionT>(opName, wires, false, std::vector<PrecisionT>{PrecisionT(0)}, {});
        double s = 0;
        check(lgh_apply_generator(sv_, b.data(), &s));
        return static_cast<PrecisionT>(s);
    }

    // ---- BLAS-like (cuda_helpers.hpp:456-560)
    ComplexT innerProduct(const StateVectorHIPManaged &other) const {
        double re = 0, im = 0;
        check(lgh_inner(sv_, other.sv_, &re, &im));
        return {static_cast<PrecisionT>(re), static_cast<PrecisionT>(im)};
    }
    void axpy(ComplexT alpha, const StateVectorHIPManaged &x) {
        check(lgh_axpy(sv_, alpha.real(), alpha.imag(), x.sv_));
    }
    void scale(ComplexT alpha) { check(lgh_scale(sv_, alpha.real(), alpha.imag())); }

    // ---- measurements (SVCM:702-1148)
    static std::vector<ComplexT> namedObservableMatrix(const std::string &name, std::size_t nwires) {
        using C = ComplexT;
        if (nwires != 1) return {};
        const PrecisionT r = PrecisionT(1) / std::sqrt(PrecisionT(2));
        if (name == "PauliX") return {C(0), C(1), C(1), C(0)};
        if (name == "PauliY") return {C(0), C(0, -1), C(0, 1), C(0)};
        if (name == "PauliZ") return {C(1), C(0), C(0), C(-1)};
        if (name == "Hadamard") return {C(r), C(r), C(r), C(-r)};
        if (name == "Identity") return {C(1), C(0), C(0), C(1)};
        return {};
    }
    // expval(obsName, wires, params, matrix): the matrix, when given, is the
    // observable; otherwise the name selects a built-in one.  Unlike the
    // reference's (name, param0)-keyed cache, the matrix content is always used.
    ComplexT expval(const std::string &obsName, const std::vector<std::size_t> &wires,
                    const std::vector<PrecisionT> &params = {}, const std::vector<ComplexT> &matrix = {}) {
        (void)params;
        std::vector<ComplexT> m = matrix.empty() ? namedObservableMatrix(obsName, wires.size()) : matrix;
        if (m.empty()) throw LightningException("Currently unsupported observable: " + obsName);
        return expval(wires, m);
    }
    ComplexT expval(const std::vector<std::size_t> &wires, const std::vector<ComplexT> &matrix) {
        if (matrix.empty()) throw LightningException("Currently unsupported observable");
        PL_HIP_ABORT_IF(matrix.size() != (std::size_t(1) << (2 * wires.size())),
                        "Observable matrix does not match the number of wires");
        std::vector<int32_t> w(wires.begin(), wires.end());
        std::vector<double> m(2 * matrix.size());
        for (std::size_t i = 0; i < matrix.size(); ++i) {
            m[2 * i] = matrix[i].real();
            m[2 * i + 1] = matrix[i].imag();
        }
        double re = 0, im = 0;
        check(lgh_expval_matrix(sv_, w.data(), static_cast<int32_t>(w.size()), m.data(), &re, &im));
        return {static_cast<PrecisionT>(re), static_cast<PrecisionT>(im)};
    }
    template <class IndexT>
    PrecisionT getExpectationValueOnSparseSpMV(const IndexT *csrOffsets, IndexT csrOffsets_size,
                                               const IndexT *columns, const ComplexT *values, IndexT numNNZ) {
        const int64_t rows = static_cast<int64_t>(csrOffsets_size) - 1;
        std::vector<int64_t> ptr(csrOffsets, csrOffsets + csrOffsets_size);
        std::vector<int64_t> col(columns, columns + numNNZ);
        std::vector<double> val(2 * static_cast<std::size_t>(numNNZ));
        for (IndexT i = 0; i < numNNZ; ++i) {
            val[2 * i] = values[i].real();
            val[2 * i + 1] = values[i].imag();
        }
        lgh_csr *csr = nullptr;
        check(lgh_csr_create(&csr, getDeviceID(), kDtype, rows, ptr.data(), col.data(), val.data(),
                             static_cast<int64_t>(numNNZ)));
        double re = 0, im = 0;
        lgh_status st = lgh_expval_csr(sv_, csr, &re, &im);
        (void)lgh_csr_destroy(csr);
        check(st);
        return static_cast<PrecisionT>(re);
    }
    PrecisionT expvalCSR(const lgh_csr *csr) const {
        double re = 0, im = 0;
        check(lgh_expval_csr(sv_, csr, &re, &im));
        return static_cast<PrecisionT>(re);
    }
    PrecisionT getExpectationValuePauliWords(const std::vector<std::string> &pauli_words,
                                             const std::vector<std::vector<std::size_t>> &tgts,
                                             const ComplexT *coeffs) {
        PL_HIP_ABORT_IF(pauli_words.size() != tgts.size(), "Pauli words and wires differ in length");
        std::string chars;
        std::vector<int32_t> wires;
        std::vector<int64_t> offsets{0};
        std::vector<double> c(2 * pauli_words.size());
        for (std::size_t w = 0; w < pauli_words.size(); ++w) {
            PL_HIP_ABORT_IF(pauli_words[w].size() != tgts[w].size(), "Pauli word and its wires differ in length");
            chars += pauli_words[w];
            for (auto t : tgts[w]) wires.push_back(static_cast<int32_t>(t));
            offsets.push_back(static_cast<int64_t>(chars.size()));
            c[2 * w] = coeffs[w].real();
            c[2 * w + 1] = coeffs[w].imag();
        }
        double res = 0;
        check(lgh_expval_pauli_words(sv_, static_cast<int64_t>(pauli_words.size()), chars.data(), wires.data(),
                                     offsets.data(), c.data(), nullptr, &res));
        return static_cast<PrecisionT>(res);
    }
    // probability(wires): wires[0] is the least significant output bit, as
    // the reference returns it (SVCM:931-970).
    std::vector<double> probability(const std::vector<std::size_t> &wires) {
        std::vector<int32_t> w(wires.begin(), wires.end());
        std::vector<double> out(std::size_t(1) << wires.size());
        check(lgh_probs(sv_, w.data(), static_cast<int32_t>(w.size()), out.data()));
        return out;
    }
    // generate_samples (SVCM:982-1061): (num_samples x num_qubits) bits,
    // ascending bitstring order.  The reference draws from an unseeded
    // generator; here each call uses the next seed of a per-state sequence
    // (setSeed fixes it for reproducible runs).
    std::vector<std::size_t> generate_samples(std::size_t num_samples) {
        return generate_samples(num_samples, seed_ + sample_calls_++);
    }
    std::vector<std::size_t> generate_samples(std::size_t num_samples, uint64_t seed) {
        static_assert(sizeof(std::size_t) == sizeof(uint64_t));
        std::vector<std::size_t> bits(num_samples * num_qubits_);
        check(lgh_sample_bits(sv_, num_samples, seed, reinterpret_cast<uint64_t *>(bits.data())));
        return bits;
    }
    std::vector<uint64_t> sample_indices(std::size_t num_samples, uint64_t seed) {
        std::vector<uint64_t> idx(num_samples);
        check(lgh_sample(sv_, num_samples, seed, idx.data()));
        return idx;
    }
    void setSeed(uint64_t seed) {
        seed_ = seed;
        sample_calls_ = 0;
    }

    // ---- named gates (Bindings.cpp:120-427)
#define LGH_NAMED(fn, name)                                                                               \
    void fn(const std::vector<std::size_t> &wires, bool adjoint) { applyOperation(name, wires, adjoint, {}); }
    LGH_NAMED(applyPauliX, "PauliX")
    LGH_NAMED(applyPauliY, "PauliY")
    LGH_NAMED(applyPauliZ, "PauliZ")
    LGH_NAMED(applyHadamard, "Hadamard")
    LGH_NAMED(applyS, "S")
    LGH_NAMED(applyT, "T")
    LGH_NAMED(applyCNOT, "CNOT")
    LGH_NAMED(applySWAP, "SWAP")
    LGH_NAMED(applyCSWAP, "CSWAP")
    LGH_NAMED(applyToffoli, "Toffoli")
    LGH_NAMED(applyCY, "CY")
    LGH_NAMED(applyCZ, "CZ")
#undef LGH_NAMED

  private:
    lgh_sv *sv_ = nullptr;
    DevTag tag_;
    std::size_t num_qubits_ = 0;
    uint64_t seed_ = std::random_device{}() ^ (uint64_t(std::random_device{}()) << 32);
    uint64_t sample_calls_ = 0;
};

} // namespace Pennylane::HIP
