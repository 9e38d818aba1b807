// pybind11 module `lightning_gpu_qubit_ops`: the same classes, methods and
// overloads as the reference module (pennylane_lightning_gpu/src/bindings/
// Bindings.cpp:67-877, :1727-1787), so that the reference's _serialize.py
// and device class can import it unchanged.  Every method forwards to the
// host classes in csrc/host, which call the C ABI of liblgh.so.
//
// The GIL is released around calls that run device work.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "AdjointJacobianHIP.hpp"
#include "ObservablesHIP.hpp"
#include "StateVectorHIPManaged.hpp"

namespace py = pybind11;
using namespace Pennylane::HIP;

namespace {

int g_current_device = 0;

// DevicePool (pennylane_lightning_gpu/src/util/DevicePool.hpp:15-128)
class DevicePool {
  public:
    DevicePool() {
        for (std::size_t i = 0; i < getGPUCount(); ++i) available_.insert(int(i));
    }
    const std::set<int> &getActiveDevices() const { return active_; }
    bool isActive(int id) const { return active_.count(id) > 0; }
    bool isInactive(int id) const { return available_.count(id) > 0; }
    int acquireDevice() {
        PL_HIP_ABORT_IF(available_.empty(), "No devices available");
        int d = *available_.begin();
        available_.erase(d);
        active_.insert(d);
        return d;
    }
    void releaseDevice(int id) {
        if (active_.erase(id)) available_.insert(id);
    }
    void syncDevice() { check(lgh_device_synchronize(g_current_device)); }
    static std::size_t getTotalDevices() { return getGPUCount(); }
    static std::vector<std::string> getDeviceUIDs() {
        std::vector<std::string> out;
        for (std::size_t i = 0; i < getGPUCount(); ++i) {
            char name[64];
            check(lgh_device_arch(int(i), name, sizeof name));
            out.push_back(std::string(name) + ":" + std::to_string(i));
        }
        return out;
    }
    static void setDeviceIdx(int id) { g_current_device = id; }

  private:
    std::set<int> available_, active_;
};

const std::vector<std::string> kFixedGates = {"PauliX", "PauliY", "PauliZ", "Hadamard", "S",    "T", "CNOT",
                                              "SWAP",   "CSWAP",  "Toffoli", "CY",      "CZ"};
const std::vector<std::string> kParamGates = {"PhaseShift",
                                              "ControlledPhaseShift",
                                              "RX",
                                              "RY",
                                              "RZ",
                                              "Rot",
                                              "CRX",
                                              "CRY",
                                              "CRZ",
                                              "CRot",
                                              "IsingXX",
                                              "IsingYY",
                                              "IsingZZ",
                                              "SingleExcitation",
                                              "SingleExcitationMinus",
                                              "SingleExcitationPlus",
                                              "DoubleExcitation",
                                              "DoubleExcitationMinus",
                                              "DoubleExcitationPlus",
                                              "MultiRZ"};

template <class T> std::vector<std::complex<T>> to_vec(const py::array_t<std::complex<T>, py::array::c_style | py::array::forcecast> &a) {
    const auto b = a.request();
    const auto *p = static_cast<const std::complex<T> *>(b.ptr);
    return std::vector<std::complex<T>>(p, p + b.size);
}

// GenerateSamples returns a (shots, num_wires) array of the state's sample
// bits (Bindings.cpp:536-553), so num_wires must be the state's qubit count:
// any other value would describe a buffer of a different size.
template <class SV> void check_sample_wires(const SV &sv, std::size_t num_wires) {
    if (num_wires != sv.getNumQubits())
        throw LightningException("GenerateSamples: num_wires (" + std::to_string(num_wires) +
                                 ") must equal the number of qubits of the state (" +
                                 std::to_string(sv.getNumQubits()) + ")");
}

py::array_t<std::size_t> samples_array(const std::vector<std::size_t> &bits, std::size_t shots,
                                       std::size_t num_wires) {
    if (bits.size() != shots * num_wires)
        throw LightningException("GenerateSamples: sample buffer does not match (shots, num_wires)");
    py::array_t<std::size_t> out({shots, num_wires});
    if (!bits.empty()) std::memcpy(out.mutable_data(), bits.data(), bits.size() * sizeof(std::size_t));
    return out;
}

template <class PrecisionT, class ParamT> void bind_precision(py::module_ &m) {
    using SV = StateVectorHIPManaged<PrecisionT>;
    using CT = std::complex<PrecisionT>;
    using np_arr_r = py::array_t<ParamT, py::array::c_style | py::array::forcecast>;
    using np_arr_c = py::array_t<std::complex<ParamT>, py::array::c_style | py::array::forcecast>;
    using IdxT = std::conditional_t<std::is_same_v<ParamT, float>, int32_t, int64_t>;
    using np_arr_idx = py::array_t<IdxT, py::array::c_style | py::array::forcecast>;
    const std::string bits = std::to_string(sizeof(CT) * 8);

    auto sv_cls = py::class_<SV>(m, ("LightningGPU_C" + bits).c_str());
    sv_cls.def(py::init([](std::size_t n) { return new SV(n, DevTag(g_current_device)); }))
        .def(py::init<std::size_t, DevTag>())
        .def(py::init<const SV &>())
        .def(py::init([](const np_arr_c &arr) {
            const auto b = arr.request();
            return new SV(static_cast<const CT *>(b.ptr), static_cast<std::size_t>(b.size), DevTag(g_current_device));
        }))
        .def(
            "setBasisState",
            [](SV &sv, std::size_t index, bool use_async) { sv.setBasisState(CT(1, 0), index, use_async); },
            "Create Basis State on GPU.")
        .def(
            "setStateVector",
            [](SV &sv, const np_arr_idx &indices, const np_arr_c &state, bool use_async) {
                const auto bi = indices.request();
                const auto bs = state.request();
                PL_HIP_ABORT_IF(bi.size != bs.size, "indices and values differ in length");
                sv.template setStateVector<IdxT>(static_cast<IdxT>(bi.size), static_cast<const CT *>(bs.ptr),
                                                 static_cast<const IdxT *>(bi.ptr), use_async);
            },
            "Set State Vector on GPU with values and their corresponding indices for the state vector on device")
        .def(
            "Identity", [](SV &, const std::vector<std::size_t> &, bool, const std::vector<ParamT> &) {},
            "Apply the Identity gate.");
    for (const auto &name : kFixedGates) {
        sv_cls.def(
            name.c_str(),
            [name](SV &sv, const std::vector<std::size_t> &wires, bool adjoint, const std::vector<ParamT> &) {
                sv.applyOperation(name, wires, adjoint, {});
            },
            ("Apply the " + name + " gate.").c_str());
    }
    for (const auto &name : kParamGates) {
        sv_cls.def(
            name.c_str(),
            [name](SV &sv, const std::vector<std::size_t> &wires, bool adjoint, const std::vector<ParamT> &params) {
                std::vector<PrecisionT> p(params.begin(), params.end());
                if (name != "Rot" && name != "CRot" && p.size() > 1) p.resize(1);
                sv.applyOperation(name, wires, adjoint, p);
            },
            ("Apply the " + name + " gate.").c_str());
    }
    sv_cls
        .def("apply",
             py::overload_cast<const std::vector<std::string> &, const std::vector<std::vector<std::size_t>> &,
                               const std::vector<bool> &, const std::vector<std::vector<PrecisionT>> &>(
                 &SV::applyOperation))
        .def("apply", py::overload_cast<const std::vector<std::string> &,
                                        const std::vector<std::vector<std::size_t>> &, const std::vector<bool> &>(
                          &SV::applyOperation))
        .def("apply", [](SV &sv, const std::string &name, const std::vector<std::size_t> &wires, bool inv,
                         const std::vector<PrecisionT> &params, const std::vector<CT> &matrix) {
            sv.applyOperation_std(name, wires, inv, params, matrix);
        })
        .def(
            "ExpectationValue",
            [](SV &sv, const std::string &obsName, const std::vector<std::size_t> &wires,
               const std::vector<ParamT> &params, const np_arr_c &gate_matrix) {
                std::vector<PrecisionT> p(params.begin(), params.end());
                auto mat = to_vec<ParamT>(gate_matrix);
                return sv.expval(obsName, wires, p, std::vector<CT>(mat.begin(), mat.end())).real();
            },
            "Calculate the expectation value of the given observable.")
        .def(
            "ExpectationValue",
            [](SV &sv, const std::vector<std::string> &obsNames, const std::vector<std::size_t> &wires,
               const std::vector<std::vector<ParamT>> &, const np_arr_c &gate_matrix) {
                std::string concat{"#"};
                for (const auto &s : obsNames) concat += s;
                auto mat = to_vec<ParamT>(gate_matrix);
                return sv.expval(concat, wires, {}, std::vector<CT>(mat.begin(), mat.end())).real();
            },
            "Calculate the expectation value of the given observable.")
        .def(
            "ExpectationValue",
            [](SV &sv, const std::vector<std::size_t> &wires, const np_arr_c &gate_matrix) {
                auto mat = to_vec<ParamT>(gate_matrix);
                return sv.expval(wires, std::vector<CT>(mat.begin(), mat.end())).real();
            },
            "Calculate the expectation value of the Hamiltonian observable.")
        .def(
            "ExpectationValue",
            [](SV &sv, const np_arr_idx &csrOffsets, const np_arr_idx &columns, const np_arr_c &values) {
                const auto bo = csrOffsets.request(), bc = columns.request(), bv = values.request();
                return sv.template getExpectationValueOnSparseSpMV<IdxT>(
                    static_cast<const IdxT *>(bo.ptr), static_cast<IdxT>(bo.size), static_cast<const IdxT *>(bc.ptr),
                    static_cast<const CT *>(bv.ptr), static_cast<IdxT>(bv.size));
            },
            "Calculate the expectation value of a sparse Hamiltonian.")
        .def(
            "ExpectationValue",
            [](SV &sv, const std::vector<std::string> &pauli_words,
               const std::vector<std::vector<std::size_t>> &target_wires, const np_arr_c &coeffs) {
                const auto b = coeffs.request();
                return sv.getExpectationValuePauliWords(pauli_words, target_wires, static_cast<const CT *>(b.ptr));
            },
            "Calculate the expectation value of a Hamiltonian composed solely from sums of Pauli-words")
        .def(
            "Probability",
            [](SV &sv, const std::vector<std::size_t> &wires) {
                auto p = sv.probability(wires);
                return py::array_t<ParamT>(py::cast(p));
            },
            "Calculate the probabilities for given wires. Results returned in Col-major order.")
        .def("GenerateSamples",
             [](SV &sv, std::size_t num_wires, std::size_t num_shots) {
                 check_sample_wires(sv, num_wires);
                 return samples_array(sv.generate_samples(num_shots), num_shots, num_wires);
             })
        .def("GenerateSamples",
             [](SV &sv, std::size_t num_wires, std::size_t num_shots, uint64_t seed) {
                 check_sample_wires(sv, num_wires);
                 return samples_array(sv.generate_samples(num_shots, seed), num_shots, num_wires);
             })
        .def("GenerateSampleIndices",
             [](SV &sv, std::size_t num_shots, uint64_t seed) {
                 auto r = sv.sample_indices(num_shots, seed);
                 py::array_t<uint64_t> out(num_shots);
                 std::memcpy(out.mutable_data(), r.data(), r.size() * sizeof(uint64_t));
                 return out;
             })
        .def("setSeed", &SV::setSeed)
        .def("applyGenerator", &SV::applyGenerator)
        .def("InnerProduct", [](const SV &a, const SV &b) { return a.innerProduct(b); })
        .def(
            "DeviceToDevice", [](SV &sv, const SV &other, bool async) { sv.updateData(other, async); },
            "Synchronize data from another GPU device to current device.")
        .def(
            "DeviceToHost",
            [](const SV &sv, np_arr_c &cpu_sv, bool) {
                auto b = cpu_sv.request();
                if (b.size) sv.CopyGpuDataToHost(static_cast<CT *>(b.ptr), static_cast<std::size_t>(b.size));
            },
            "Synchronize data from the GPU device to host.")
        .def(
            "HostToDevice",
            [](SV &sv, const np_arr_c &cpu_sv, bool async) {
                const auto b = cpu_sv.request();
                const auto len = static_cast<std::size_t>(b.shape.empty() ? 0 : b.size);
                if (len) sv.CopyHostDataToGpu(static_cast<const CT *>(b.ptr), len, async);
            },
            "Synchronize data from the host device to GPU.")
        .def("GetNumGPUs", [](const SV &) { return getGPUCount(); }, "Get the number of available GPUs.")
        .def("getCurrentGPU", [](const SV &sv) { return sv.getDeviceID(); },
             "Get the GPU index for the statevector data.")
        .def("numQubits", &SV::getNumQubits)
        .def("dataLength", &SV::getLength)
        .def("resetGPU", [](SV &sv, bool async) { sv.initSV(async); }, py::arg("async") = false);

    // ---- observables (Bindings.cpp:616-767)
    using Obs = ObservableHIP<PrecisionT>;
    using ObsPtr = std::shared_ptr<Obs>;
    py::class_<Obs, ObsPtr>(m, ("ObservableGPU_C" + bits).c_str(), py::module_local());

    auto eq = [](auto *tag) {
        using O = std::remove_pointer_t<decltype(tag)>;
        return [](const O &self, py::handle other) -> bool {
            if (!py::isinstance<O>(other)) return false;
            return self == other.cast<const O &>();
        };
    };

    using Named = NamedObsHIP<PrecisionT>;
    py::class_<Named, std::shared_ptr<Named>, Obs>(m, ("NamedObsGPU_C" + bits).c_str(), py::module_local())
        .def(py::init([](const std::string &name, const std::vector<std::size_t> &wires) {
            return Named(name, wires);
        }))
        .def("__repr__", &Named::getObsName)
        .def("get_wires", &Named::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Named *>(nullptr)), "Compare two observables");

    using Herm = HermitianObsHIP<PrecisionT>;
    py::class_<Herm, std::shared_ptr<Herm>, Obs>(m, ("HermitianObsGPU_C" + bits).c_str(), py::module_local())
        .def(py::init([](const np_arr_c &matrix, const std::vector<std::size_t> &wires) {
            auto v = to_vec<ParamT>(matrix);
            return Herm(std::vector<CT>(v.begin(), v.end()), wires);
        }))
        .def("__repr__", &Herm::getObsName)
        .def("get_wires", &Herm::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Herm *>(nullptr)), "Compare two observables");

    using Tensor = TensorProdObsHIP<PrecisionT>;
    py::class_<Tensor, std::shared_ptr<Tensor>, Obs>(m, ("TensorProdObsGPU_C" + bits).c_str(), py::module_local())
        .def(py::init([](const std::vector<ObsPtr> &obs) { return Tensor(obs); }))
        .def("__repr__", &Tensor::getObsName)
        .def("get_wires", &Tensor::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Tensor *>(nullptr)), "Compare two observables");

    using Ham = HamiltonianHIP<PrecisionT>;
    py::class_<Ham, std::shared_ptr<Ham>, Obs>(m, ("HamiltonianGPU_C" + bits).c_str(), py::module_local())
        .def(py::init([](const np_arr_r &coeffs, const std::vector<ObsPtr> &obs) {
            const auto b = coeffs.request();
            const auto *p = static_cast<const ParamT *>(b.ptr);
            return Ham(std::vector<PrecisionT>(p, p + b.size), obs);
        }))
        .def("__repr__", &Ham::getObsName)
        .def("get_wires", &Ham::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Ham *>(nullptr)), "Compare two observables");

    using Sparse = SparseHamiltonianHIP<PrecisionT>;
    py::class_<Sparse, std::shared_ptr<Sparse>, Obs>(m, ("SparseHamiltonianGPU_C" + bits).c_str(),
                                                     py::module_local())
        .def(py::init([](const np_arr_c &data, const np_arr_idx &indices, const np_arr_idx &offsets,
                         const std::vector<std::size_t> &wires) {
            auto d = to_vec<ParamT>(data);
            const auto bi = indices.request(), bo = offsets.request();
            const auto *ip = static_cast<const IdxT *>(bi.ptr);
            const auto *op = static_cast<const IdxT *>(bo.ptr);
            return Sparse(std::vector<CT>(d.begin(), d.end()), std::vector<IdxT>(ip, ip + bi.size),
                          std::vector<IdxT>(op, op + bo.size), wires);
        }))
        .def("__repr__", &Sparse::getObsName)
        .def("get_wires", &Sparse::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Sparse *>(nullptr)), "Compare two observables")
        .def("expval", [](const Sparse &h, const SV &sv) { return h.expval(sv); },
             "Expectation value with the CSR kept resident on the device.");

    // ---- operations and adjoint (Bindings.cpp:773-876)
    using Ops = OpsData<PrecisionT>;
    py::class_<Ops>(m, ("OpsStructGPU_C" + bits).c_str(), py::module_local())
        .def(py::init<const std::vector<std::string> &, const std::vector<std::vector<PrecisionT>> &,
                      const std::vector<std::vector<std::size_t>> &, const std::vector<bool> &,
                      const std::vector<std::vector<CT>> &>())
        .def("__repr__", [](const Ops &ops) {
            std::ostringstream s;
            for (std::size_t i = 0; i < ops.getSize(); ++i) {
                s << "{'name': " << ops.getOpsName()[i] << ", 'params': " << wiresToString(ops.getOpsParams()[i])
                  << ", 'inv': " << (ops.getOpsInverses()[i] ? 1 : 0) << "}";
                if (i + 1 < ops.getSize()) s << ",";
            }
            return "Operations: [" + s.str() + "]";
        });

    using Adj = AdjointJacobianHIP<PrecisionT>;
    py::class_<Adj>(m, ("AdjointJacobianGPU_C" + bits).c_str(), py::module_local())
        .def(py::init<>())
        .def("create_ops_list",
             [](Adj &, const std::vector<std::string> &names, const std::vector<np_arr_r> &params,
                const std::vector<std::vector<std::size_t>> &wires, const std::vector<bool> &inverses,
                const std::vector<np_arr_c> &matrices) {
                 std::vector<std::vector<PrecisionT>> p(params.size());
                 std::vector<std::vector<CT>> mats(matrices.size());
                 for (std::size_t i = 0; i < params.size(); ++i) {
                     const auto b = params[i].request();
                     const auto *ptr = static_cast<const ParamT *>(b.ptr);
                     p[i].assign(ptr, ptr + b.size);
                 }
                 for (std::size_t i = 0; i < matrices.size(); ++i) {
                     auto v = to_vec<ParamT>(matrices[i]);
                     mats[i].assign(v.begin(), v.end());
                 }
                 return Ops(names, p, wires, inverses, mats);
             })
        .def("adjoint_jacobian",
             [](Adj &adj, const SV &sv, const std::vector<ObsPtr> &obs, const Ops &ops,
                const std::vector<std::size_t> &tp) {
                 std::vector<std::vector<PrecisionT>> jac;
                 {
                     py::gil_scoped_release release;
                     jac = adj.adjointJacobian(sv, obs, ops, tp, false);
                 }
                 return py::array_t<ParamT>(py::cast(jac));
             })
        .def("adjoint_jacobian_batched",
             [](Adj &adj, const SV &sv, const std::vector<ObsPtr> &obs, const Ops &ops,
                const std::vector<std::size_t> &tp) {
                 std::vector<std::vector<PrecisionT>> jac;
                 {
                     py::gil_scoped_release release;
                     jac = adj.batchAdjointJacobian(sv, obs, ops, tp, false);
                 }
                 return py::array_t<ParamT>(py::cast(jac));
             });
}

} // namespace

PYBIND11_MODULE(lightning_gpu_qubit_ops, m) {
    py::options options;
    options.disable_function_signatures();
    py::register_exception<LightningException>(m, "PLException");

    m.def("device_reset", [] { check(lgh_device_reset()); }, "Reset all GPU devices and contexts.");
    m.def("allToAllAccess", [] {}, "Peer access between devices (a no-op on one device).");
    m.def(
        "is_gpu_supported",
        [](int device_number) {
            int s = 0;
            check(lgh_device_supported(device_number, &s));
            return s != 0;
        },
        py::arg("device_number") = 0,
        "Checks if the given GPU device meets the minimum architecture support for the PennyLane-Lightning-GPU "
        "device.");
    m.def(
        "get_gpu_arch",
        [](int device_number) {
            char name[64];
            check(lgh_device_arch(device_number, name, sizeof name));
            // "gfxMmm": major = M, minor = next digit (gfx950 -> (9, 5))
            std::string s(name);
            int major = 0, minor = 0;
            if (s.size() >= 5 && s.rfind("gfx", 0) == 0) {
                major = s[3] - '0';
                minor = s[4] - '0';
            }
            return std::make_pair(major, minor);
        },
        py::arg("device_number") = 0, "Returns the given GPU major and minor GPU support.");
    m.def("get_gpu_arch_name", [](int device_number) {
        char name[64];
        check(lgh_device_arch(device_number, name, sizeof name));
        return std::string(name);
    }, py::arg("device_number") = 0);

    py::class_<DevicePool>(m, "DevPool")
        .def(py::init<>())
        .def("getActiveDevices", &DevicePool::getActiveDevices)
        .def("isActive", &DevicePool::isActive)
        .def("isInactive", &DevicePool::isInactive)
        .def("acquireDevice", &DevicePool::acquireDevice)
        .def("releaseDevice", &DevicePool::releaseDevice)
        .def("syncDevice", &DevicePool::syncDevice)
        .def_static("getTotalDevices", &DevicePool::getTotalDevices)
        .def_static("getDeviceUIDs", &DevicePool::getDeviceUIDs)
        .def_static("setDeviceID", &DevicePool::setDeviceIdx);

    py::class_<DevTag>(m, "DevTag")
        .def(py::init<>())
        .def(py::init<int>())
        .def(py::init([](int device_id, std::uintptr_t stream) {
            return new DevTag(device_id, reinterpret_cast<void *>(stream));
        }))
        .def(py::init<const DevTag &>())
        .def("getDeviceID", &DevTag::getDeviceID)
        .def("getStreamID", [](DevTag &t) { return reinterpret_cast<std::uintptr_t>(t.getStreamID()); })
        .def("refresh", &DevTag::refresh);

    bind_precision<float, float>(m);
    bind_precision<double, double>(m);
}
