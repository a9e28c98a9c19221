// pybind11 module psi4_amd._core: the JK object over the C ABI.
//
// Mirrors psi4's Python surface of JK (psi4/src/export_fock.cc:49-97):
// build_JK, name, memory_estimate, initialize, set_print, set_cutoff,
// set_memory, set_omp_nthread, set_do_J/K/wK, set_omega(_alpha/_beta),
// set_wcombine, compute, finalize, C_clear, C_add, C_left_add, C_right_add,
// J, K, wK, D, print_header.  Matrices are numpy float64 arrays (psi4 uses
// core.Matrix).  Errors surface as RuntimeError, as psi4's PsiException does.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdio>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "psi4amd_dfjk.h"

namespace py = pybind11;
using Arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc) {
    if (rc != DFJK_OK) throw std::runtime_error(dfjk_last_error());
}

// Flat basis arrays pulled from a psi4_amd.basisset.BasisSet-like object.
struct BasisArrays {
    py::array_t<int, py::array::c_style | py::array::forcecast> l, nprim;
    Arr center, exps, coefs;
    int puream = 1;
    p4a_basis desc{};
    explicit BasisArrays(const py::object& b) {
        using IArr = py::array_t<int, py::array::c_style | py::array::forcecast>;
        l = py::cast<IArr>(b.attr("l"));
        nprim = py::cast<IArr>(b.attr("nprim"));
        center = py::cast<Arr>(b.attr("center"));
        exps = py::cast<Arr>(b.attr("exps"));
        coefs = py::cast<Arr>(b.attr("coefs"));
        puream = py::cast<bool>(b.attr("puream")) ? 1 : 0;
        desc.nshell = static_cast<int>(l.size());
        desc.l = l.data();
        desc.nprim = nprim.data();
        desc.center = center.data();
        desc.exps = exps.data();
        desc.coefs = coefs.data();
        desc.puream = puream;
    }
};

class JK {
   public:
    JK(dfjk_handle* h, int nbf, std::string name) : h_(h), nbf_(nbf), name_(std::move(name)) {}
    ~JK() {
        if (h_) dfjk_destroy(h_);
    }

    static std::shared_ptr<JK> build(const py::object& primary, const py::object& aux, bool do_wK, size_t doubles) {
        BasisArrays p(primary), a(aux);
        dfjk_handle* h = nullptr;
        check(dfjk_create(&h, &p.desc, &a.desc));
        int nbf = 0;
        check(dfjk_get_info(h, &nbf, nullptr, nullptr, nullptr));
        auto jk = std::make_shared<JK>(h, nbf, "MemDFJK");
        jk->nshell_ = p.desc.nshell;
        jk->set_do_wK(do_wK);
        jk->set_memory(doubles);
        return jk;
    }

    static std::shared_ptr<JK> build_tensor(int nbf, int naux) {
        dfjk_handle* h = nullptr;
        check(dfjk_create_tensor(&h, nbf, naux));
        return std::make_shared<JK>(h, nbf, "MemDFJK");
    }

    std::string name() const { return name_; }
    size_t memory_estimate() {
        size_t d = 0;
        check(dfjk_memory_estimate(h_, &d));
        return d;
    }
    size_t device_footprint() {
        size_t d = 0;
        check(dfjk_device_footprint(h_, &d));
        return d;
    }
    // Schwarz function-pair mask (nbf x nbf, 0/1) of the engine's screening
    py::array_t<uint8_t> pair_mask() {
        py::array_t<uint8_t> m({static_cast<py::ssize_t>(nbf_), static_cast<py::ssize_t>(nbf_)});
        check(dfjk_get_pair_mask(h_, m.mutable_data(), nullptr));
        return m;
    }
    // screening integrals: (|(mn|mn)| [nbf][nbf], shell maxima [nshell][nshell], max, on_device)
    py::tuple schwarz_values() {
        double mx = 0.0;
        int dev = 0;
        check(dfjk_get_schwarz(h_, nullptr, nullptr, &mx, &dev));
        const py::ssize_t n = static_cast<py::ssize_t>(nbf_), ns = static_cast<py::ssize_t>(nshell_);
        py::array_t<double> f({n, n}), sh({ns, ns});
        check(dfjk_get_schwarz(h_, f.mutable_data(), sh.mutable_data(), nullptr, nullptr));
        return py::make_tuple(f, sh, mx, dev != 0);
    }
    // declared structure of non-symmetric pairs: "DETECT", "GENERAL" or "CROSS"
    void set_pair_form(const std::string& f) {
        int v;
        if (f == "DETECT")
            v = DFJK_FORM_DETECT;
        else if (f == "GENERAL")
            v = DFJK_FORM_GENERAL;
        else if (f == "CROSS")
            v = DFJK_FORM_CROSS;
        else
            throw std::runtime_error("JK: pair form must be DETECT, GENERAL or CROSS");
        check(dfjk_set_pair_form(h_, v));
    }
    void initialize() {
        check(dfjk_initialize(h_));
        initialized_ = true;
    }
    void set_print(int p) { print_ = p; }
    void set_cutoff(double c) { check(dfjk_set_cutoff(h_, c)); cutoff_ = c; }
    void set_memory(size_t d) { check(dfjk_set_memory(h_, d)); memory_ = d; }
    void set_omp_nthread(int n) { check(dfjk_set_omp_nthread(h_, n)); nthread_ = n; }
    void set_do_J(bool v) { check(dfjk_set_do_J(h_, v)); do_J_ = v; }
    void set_do_K(bool v) { check(dfjk_set_do_K(h_, v)); do_K_ = v; }
    void set_do_wK(bool v) { check(dfjk_set_do_wK(h_, v)); do_wK_ = v; }
    // SCF_SUBTYPE: AUTO, INCORE, OUT_OF_CORE (jk.h:1039 set_subalgo)
    void set_subalgo(const std::string& v) {
        int a;
        if (v == "AUTO")
            a = DFJK_SUBALGO_AUTO;
        else if (v == "INCORE")
            a = DFJK_SUBALGO_INCORE;
        else if (v == "OUT_OF_CORE")
            a = DFJK_SUBALGO_OUT_OF_CORE;
        else
            throw std::runtime_error("JK: SCF_SUBTYPE must be AUTO, INCORE or OUT_OF_CORE");
        check(dfjk_set_subalgorithm(h_, a));
    }
    void set_omega(double w) { check(dfjk_set_omega(h_, w)); omega_ = w; }
    double get_omega() const { return omega_; }
    // combined exchange (MemDFJK.cc:83-90, 140-154): with do_wK, wK = alpha K + beta wK(omega), K not built
    void set_wcombine(bool w) { check(dfjk_set_wcombine(h_, w)); wcombine_ = w; }
    bool get_wcombine() const { return wcombine_; }
    void set_omega_alpha(double a) { check(dfjk_set_omega_alpha(h_, a)); omega_alpha_ = a; }
    double get_omega_alpha() const { return omega_alpha_; }
    void set_omega_beta(double b) { check(dfjk_set_omega_beta(h_, b)); omega_beta_ = b; }
    double get_omega_beta() const { return omega_beta_; }
    void set_condition(double c) { check(dfjk_set_condition(h_, c)); condition_ = c; }
    void set_precision(int p) { check(dfjk_set_precision(h_, p)); precision_ = p; }
    void set_fitting(int f) { check(dfjk_set_fitting(h_, f)); }
    void set_profiling(bool on) { check(dfjk_set_profiling(h_, on)); }
    void set_aux_range(int q0, int q1) {
        check(dfjk_set_aux_range(h_, q0, q1));
        aux_q0_ = q0;
        aux_q1_ = q1;
    }
    void set_accumulate(bool a) { accumulate_ = a; }

    py::list& C_left() { return Cl_; }
    py::list& C_right() { return Cr_; }
    void C_clear() {
        Cl_ = py::list();
        Cr_ = py::list();
    }
    void C_add(const Arr& C) {
        Cl_.append(C);
        Cr_.append(C);
    }
    void C_left_add(const Arr& C) { Cl_.append(C); }
    void C_right_add(const Arr& C) { Cr_.append(C); }

    // JK::compute (libfock/jk.cc:595-681): lr_symmetric if C_right is empty.
    void compute() {
        const bool lr_symmetric = py::len(Cl_) > 0 && py::len(Cr_) == 0;
        const size_t n = py::len(Cl_);
        if (!lr_symmetric && py::len(Cr_) != n) throw std::runtime_error("JK: C_left/C_right irrep mismatch!");
        std::vector<Arr> cl, cr;
        std::vector<int> ncol;
        for (size_t i = 0; i < n; ++i) {
            Arr L = Arr::ensure(Cl_[i]);
            if (!L || L.ndim() != 2 || L.shape(0) != nbf_) throw std::runtime_error("JK: C_left must be nbf x ncol");
            cl.push_back(L);
            ncol.push_back(static_cast<int>(L.shape(1)));
            if (!lr_symmetric) {
                Arr R = Arr::ensure(Cr_[i]);
                if (!R || R.ndim() != 2 || R.shape(0) != nbf_) throw std::runtime_error("JK: C_right must be nbf x ncol");
                if (R.shape(1) != L.shape(1))
                    throw std::runtime_error("JK: C_left/C_right MO zip index size mismatch!");
                cr.push_back(R);
            }
        }
        if (!initialized_) throw std::runtime_error("JK: initialize() must be called before compute()");
        // outputs are (re)allocated unless accumulating into existing ones of the right count
        if (!accumulate_ || J_.size() != n || (do_wK_ && wK_.size() != n)) {
            D_.clear();
            J_.clear();
            K_.clear();
            wK_.clear();
            for (size_t i = 0; i < n; ++i) {
                D_.push_back(Arr({nbf_, nbf_}));
                J_.push_back(Arr({nbf_, nbf_}));
                K_.push_back(Arr({nbf_, nbf_}));
                std::fill(J_.back().mutable_data(), J_.back().mutable_data() + nbf_ * nbf_, 0.0);
                std::fill(K_.back().mutable_data(), K_.back().mutable_data() + nbf_ * nbf_, 0.0);
                if (do_wK_) {
                    wK_.push_back(Arr({nbf_, nbf_}));
                    std::fill(wK_.back().mutable_data(), wK_.back().mutable_data() + nbf_ * nbf_, 0.0);
                }
            }
        }
        std::vector<const double*> pl(n), pr(n);
        std::vector<double*> pd(n), pj(n), pk(n);
        for (size_t i = 0; i < n; ++i) {
            pl[i] = cl[i].data();
            if (!lr_symmetric) pr[i] = cr[i].data();
            pd[i] = D_[i].mutable_data();
            pj[i] = J_[i].mutable_data();
            pk[i] = K_[i].mutable_data();
        }
        {
            py::gil_scoped_release nogil;
            check(dfjk_compute(h_, static_cast<int>(n), pl.data(), lr_symmetric ? nullptr : pr.data(), ncol.data(),
                               pd.data(), do_J_ ? pj.data() : nullptr, do_K_ ? pk.data() : nullptr,
                               accumulate_ ? 1 : 0));
            if (do_wK_) {
                std::vector<double*> pw(n);
                for (size_t i = 0; i < n; ++i) pw[i] = wK_[i].mutable_data();
                check(dfjk_compute_wK(h_, static_cast<int>(n), pl.data(), lr_symmetric ? nullptr : pr.data(),
                                      ncol.data(), pw.data(), accumulate_ ? 1 : 0));
            }
        }
    }

    // Device-pointer variant (torch tensors): pointers as integers.
    void compute_device(const std::vector<uintptr_t>& Cl, const std::vector<uintptr_t>& Cr, const std::vector<int>& ncol,
                        const std::vector<uintptr_t>& D, const std::vector<uintptr_t>& J,
                        const std::vector<uintptr_t>& K, bool accumulate) {
        const size_t n = Cl.size();
        if (ncol.size() != n || (!Cr.empty() && Cr.size() != n)) throw std::runtime_error("JK: pointer count mismatch");
        auto cv = [&](const std::vector<uintptr_t>& v) {
            std::vector<double*> o(v.size());
            for (size_t i = 0; i < v.size(); ++i) o[i] = reinterpret_cast<double*>(v[i]);
            return o;
        };
        std::vector<const double*> pl(n), pr(Cr.size());
        for (size_t i = 0; i < n; ++i) pl[i] = reinterpret_cast<const double*>(Cl[i]);
        for (size_t i = 0; i < Cr.size(); ++i) pr[i] = reinterpret_cast<const double*>(Cr[i]);
        auto pd = cv(D), pj = cv(J), pk = cv(K);
        py::gil_scoped_release nogil;
        check(dfjk_compute_device(h_, static_cast<int>(n), pl.data(), Cr.empty() ? nullptr : pr.data(), ncol.data(),
                                  D.empty() ? nullptr : pd.data(), J.empty() ? nullptr : pj.data(),
                                  K.empty() ? nullptr : pk.data(), accumulate ? 1 : 0));
    }

    void set_stream(uintptr_t s) { check(dfjk_set_stream(h_, reinterpret_cast<void*>(s))); }
    void set_B_dense(const Arr& B) {
        check(dfjk_set_B_dense(h_, B.data(), 0));
    }
    void set_B_dense_device(uintptr_t B) { check(dfjk_set_B_dense(h_, reinterpret_cast<const double*>(B), 1)); }

    // Adjoint of E = tr(D(2J-K)), D = C C^T: returns (E, dE/dC, dE/dB or None).
    py::tuple energy_adjoint(const Arr& C, bool want_dEdB) {
        if (C.ndim() != 2 || C.shape(0) != nbf_) throw std::runtime_error("JK: C must be nbf x nocc");
        const int nocc = static_cast<int>(C.shape(1));
        int naux = 0;
        check(dfjk_get_info(h_, nullptr, &naux, nullptr, nullptr));
        const int q0 = aux_q1_ > aux_q0_ ? aux_q0_ : 0, q1 = aux_q1_ > aux_q0_ ? aux_q1_ : naux;
        Arr dC({static_cast<py::ssize_t>(nbf_), static_cast<py::ssize_t>(nocc)});
        py::object dB = py::none();
        double* pB = nullptr;
        if (want_dEdB) {
            Arr b({static_cast<py::ssize_t>(q1 - q0), static_cast<py::ssize_t>(nbf_), static_cast<py::ssize_t>(nbf_)});
            pB = b.mutable_data();
            dB = b;
        }
        double E = 0.0;
        double* pC = dC.mutable_data();
        {
            py::gil_scoped_release nogil;
            check(dfjk_energy_adjoint(h_, C.data(), nocc, &E, pC, pB));
        }
        return py::make_tuple(E, dC, dB);
    }
    double energy_adjoint_device(uintptr_t C, int nocc, uintptr_t dEdC, uintptr_t dEdB) {
        double E = 0.0;
        py::gil_scoped_release nogil;
        check(dfjk_energy_adjoint_device(h_, reinterpret_cast<const double*>(C), nocc, &E,
                                         reinterpret_cast<double*>(dEdC), reinterpret_cast<double*>(dEdB)));
        return E;
    }
    void fill_synthetic(uint64_t seed) { check(dfjk_fill_synthetic(h_, seed)); }
    Arr get_B_dense(int q0, int q1) {
        Arr out({static_cast<py::ssize_t>(q1 - q0), static_cast<py::ssize_t>(nbf_), static_cast<py::ssize_t>(nbf_)});
        check(dfjk_get_B_dense(h_, q0, q1, out.mutable_data()));
        return out;
    }
    py::dict timings() {
        double j, ht, kg, other;
        int nht, nkg;
        check(dfjk_get_timings(h_, &j, &ht, &kg, &other, &nht, &nkg));
        py::dict d;
        d["J"] = j;
        d["HT"] = ht;
        d["KGEMM"] = kg;
        d["other"] = other;
        d["n_ht"] = nht;
        d["n_kgemm"] = nkg;
        return d;
    }
    // the C-ABI handle behind this object, for direct calls into libpsi4amd_dfjk
    uintptr_t handle() const { return reinterpret_cast<uintptr_t>(h_); }
    py::dict info() {
        int nbf, naux, ntp;
        int64_t nsig;
        check(dfjk_get_info(h_, &nbf, &naux, &ntp, &nsig));
        py::dict d;
        d["nbf"] = nbf;
        d["naux"] = naux;
        d["ntile_pairs"] = ntp;
        d["nsig_pairs"] = nsig;
        int ooc = 0;
        check(dfjk_get_out_of_core(h_, &ooc));
        d["out_of_core"] = ooc != 0;
        double ms[6];
        check(dfjk_get_setup_timings(h_, ms, 6));
        py::dict t;
        const char* names[6] = {"screening", "metric", "metric_factor", "three_index", "fit", "attenuated"};
        for (int i = 0; i < 6; ++i) t[names[i]] = ms[i];
        d["setup_ms"] = t;
        int fitting = 0, kept = 0;
        double cond = 0.0, cutoff = 0.0;
        check(dfjk_get_fit_info(h_, &fitting, &cond, &kept));
        check(dfjk_get_pair_mask(h_, nullptr, &cutoff));
        d["fitting"] = fitting;
        d["condition"] = cond;
        d["naux_kept"] = kept;
        d["cutoff"] = cutoff;
        int q0 = 0, q1 = 0;
        check(dfjk_get_aux_range(h_, &q0, &q1));
        d["aux_range"] = py::make_tuple(q0, q1);
        return d;
    }

    py::list J() const { return to_list(J_); }
    py::list K() const { return to_list(K_); }
    py::list D() const { return to_list(D_); }
    py::list wK() const { return to_list(wK_); }
    void set_Bw_dense(Arr Bw) {
        const int naux = info_naux();
        if (Bw.ndim() != 3 || Bw.shape(0) != naux || Bw.shape(1) != nbf_ || Bw.shape(2) != nbf_)
            throw std::runtime_error("JK: Bw must be naux x nbf x nbf");
        check(dfjk_set_Bw_dense(h_, Bw.data(), 0));
    }
    Arr get_Bw_dense(int q0, int q1) {
        Arr out({static_cast<py::ssize_t>(q1 - q0), static_cast<py::ssize_t>(nbf_), static_cast<py::ssize_t>(nbf_)});
        check(dfjk_get_Bw_dense(h_, q0, q1, out.mutable_data()));
        return out;
    }
    int info_naux() {
        int naux = 0;
        check(dfjk_get_info(h_, nullptr, &naux, nullptr, nullptr));
        return naux;
    }
    void finalize() { check(dfjk_finalize(h_)); }

    // MemDFJK::print_header (libfock/MemDFJK.cc:113-132)
    std::string print_header() {
        int nbf, naux, ntp;
        int64_t nsig;
        check(dfjk_get_info(h_, &nbf, &naux, &ntp, &nsig));
        std::ostringstream os;
        char buf[256];
        os << "  ==> MemDFJK: Density-Fitted J/K Matrices <==\n\n";
        std::snprintf(buf, sizeof buf, "    J tasked:           %11s\n", do_J_ ? "Yes" : "No");
        os << buf;
        std::snprintf(buf, sizeof buf, "    K tasked:           %11s\n", do_K_ ? "Yes" : "No");
        os << buf;
        std::snprintf(buf, sizeof buf, "    wK tasked:          %11s\n", do_wK_ ? "Yes" : "No");
        os << buf;
        std::snprintf(buf, sizeof buf, "    Memory [MiB]:       %11ld\n", static_cast<long>((memory_ * 8L) / (1024L * 1024L)));
        os << buf;
        int ooc = 0;
        if (initialized_) check(dfjk_get_out_of_core(h_, &ooc));
        std::snprintf(buf, sizeof buf, "    Algorithm:          %11s\n", ooc ? "Host-stream" : "Device");
        os << buf;
        std::snprintf(buf, sizeof buf, "    Schwarz Cutoff:     %11.0E\n", cutoff_);
        os << buf;
        const double sparsity = initialized_ ? 1.0 - static_cast<double>(nsig) / (static_cast<double>(nbf) * nbf) : 0.0;
        std::snprintf(buf, sizeof buf, "    Mask sparsity (%%):  %11.4f\n", 100.0 * sparsity);
        os << buf;
        std::snprintf(buf, sizeof buf, "    Fitting Condition:  %11.0E\n\n", condition_);
        os << buf;
        std::string s = os.str();
        if (print_) py::print(s, py::arg("end") = "");
        return s;
    }

   private:
    static py::list to_list(const std::vector<Arr>& v) {
        py::list l;
        for (auto& a : v) l.append(a);
        return l;
    }
    dfjk_handle* h_;
    int nbf_;
    int nshell_ = 0;  // primary shells (basis-built handles)
    int aux_q0_ = 0, aux_q1_ = 0;
    std::string name_;
    py::list Cl_, Cr_;
    std::vector<Arr> D_, J_, K_, wK_;
    bool initialized_ = false, do_J_ = true, do_K_ = true, do_wK_ = false, wcombine_ = false, accumulate_ = false;
    int print_ = 1, nthread_ = 1, precision_ = 0;
    size_t memory_ = 256000000;
    double cutoff_ = 1e-12, condition_ = 1e-10, omega_ = 0.0, omega_alpha_ = 1.0, omega_beta_ = 0.0;
};

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "MI355X DF-JK engine (C ABI include/psi4amd_dfjk.h) exposed as psi4's JK object";
    m.attr("PREC_FP64") = static_cast<int>(DFJK_PREC_FP64);
    m.attr("PREC_FP32") = static_cast<int>(DFJK_PREC_FP32);
    m.attr("PREC_BF16X3") = static_cast<int>(DFJK_PREC_BF16X3);
    m.attr("FIT_CHOLESKY") = static_cast<int>(DFJK_FIT_CHOLESKY);
    m.attr("FIT_EIGEN") = static_cast<int>(DFJK_FIT_EIGEN);
    py::class_<JK, std::shared_ptr<JK>>(m, "JK", py::dynamic_attr())
        .def_static("build_JK", &JK::build, py::arg("basis"), py::arg("aux"), py::arg("do_wK") = false,
                    py::arg("doubles") = static_cast<size_t>(256000000))
        .def_static("build_tensor", &JK::build_tensor, py::arg("nbf"), py::arg("naux"))
        .def("name", &JK::name)
        .def("memory_estimate", &JK::memory_estimate)
        .def("device_footprint", &JK::device_footprint)
        .def("pair_mask", &JK::pair_mask)
        .def("schwarz_values", &JK::schwarz_values)
        .def("set_pair_form", &JK::set_pair_form)
        .def("initialize", &JK::initialize)
        .def("set_print", &JK::set_print)
        .def("set_cutoff", &JK::set_cutoff)
        .def("set_memory", &JK::set_memory)
        .def("set_omp_nthread", &JK::set_omp_nthread)
        .def("set_do_J", &JK::set_do_J)
        .def("set_do_K", &JK::set_do_K)
        .def("set_do_wK", &JK::set_do_wK)
        .def("set_subalgo", &JK::set_subalgo)
        .def("set_omega", &JK::set_omega, py::arg("omega"))
        .def("get_omega", &JK::get_omega)
        .def("set_wcombine", &JK::set_wcombine, py::arg("wcombine"))
        .def("get_wcombine", &JK::get_wcombine)
        .def("set_omega_alpha", &JK::set_omega_alpha, py::arg("alpha"))
        .def("get_omega_alpha", &JK::get_omega_alpha)
        .def("set_omega_beta", &JK::set_omega_beta, py::arg("beta"))
        .def("get_omega_beta", &JK::get_omega_beta)
        .def("set_condition", &JK::set_condition)
        .def("set_precision", &JK::set_precision)
        .def("set_fitting", &JK::set_fitting)
        .def("set_profiling", &JK::set_profiling)
        .def("set_aux_range", &JK::set_aux_range)
        .def("set_accumulate", &JK::set_accumulate)
        .def("set_stream", &JK::set_stream)
        .def("compute", &JK::compute)
        .def("compute_device", &JK::compute_device, py::arg("C_left"), py::arg("C_right"), py::arg("ncol"),
             py::arg("D"), py::arg("J"), py::arg("K"), py::arg("accumulate") = false)
        .def("finalize", &JK::finalize)
        .def("C_clear", &JK::C_clear)
        .def("C_add", &JK::C_add)
        .def("C_left_add", &JK::C_left_add)
        .def("C_right_add", &JK::C_right_add)
        .def("C_left", &JK::C_left, py::return_value_policy::reference_internal)
        .def("C_right", &JK::C_right, py::return_value_policy::reference_internal)
        .def("J", &JK::J)
        .def("K", &JK::K)
        .def("wK", &JK::wK)
        .def("D", &JK::D)
        .def("set_B_dense", &JK::set_B_dense)
        .def("set_B_dense_device", &JK::set_B_dense_device)
        .def("energy_adjoint", &JK::energy_adjoint, py::arg("C"), py::arg("want_dEdB") = false)
        .def("energy_adjoint_device", &JK::energy_adjoint_device, py::arg("C"), py::arg("nocc"), py::arg("dEdC"),
             py::arg("dEdB"))
        .def("fill_synthetic", &JK::fill_synthetic)
        .def("get_B_dense", &JK::get_B_dense)
        .def("set_Bw_dense", &JK::set_Bw_dense)
        .def("get_Bw_dense", &JK::get_Bw_dense)
        .def("timings", &JK::timings)
        .def("info", &JK::info)
        .def("handle", &JK::handle)
        .def("print_header", &JK::print_header);
}
