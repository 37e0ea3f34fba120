// fast_pauli_py.cpp — the Python module `_fast_pauli` over the C++ host API.
//
// Mirrors the reference's nanobind module (fast_pauli/cpp/src/fast_pauli.cpp:28-105
// and the binding headers it includes): the same classes, constructors,
// properties, methods, argument names, shapes and exception types.  States are
// 1-D (dim,) or 2-D (dim, n_states) C-contiguous arrays; complex128 is the
// default precision (other dtypes are converted, __types.hpp:5-6), and a
// complex64 array runs the complex64 / FP32-accumulation path and returns
// complex64.  Outputs are fresh arrays; the GPU writes them directly.
//
// Deliberate difference: PauliString.apply on a 1-D state honours `coeff`
// (the reference's 1-D branch drops it, __pauli_string_bindings.hpp:142,
// although its docstring promises c P|psi>).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "fast_pauli/fast_pauli.hpp"

namespace py = pybind11;
namespace fp = fast_pauli;
using namespace pybind11::literals;

using cfloat_t = std::complex<double>;
using float_type = double;

namespace
{

// A states argument: C-contiguous complex128, or complex64 for the fast path.
struct States
{
    py::array arr;
    bool c64 = false;
    std::size_t ndim = 0, dim = 0, n_states = 1;
};

States as_states(py::handle obj, char const *what = "states")
{
    States s;
    py::array a = py::array::ensure(obj);
    if (!a)
        throw py::type_error(std::string(what) + " must be array-like");
    auto const dt = a.dtype();
    if (dt.is(py::dtype::of<std::complex<float>>()))
    {
        s.c64 = true;
        if (!(a.flags() & py::array::c_style))
            throw std::invalid_argument("ndarray must be row-major (C-contiguous)");
        s.arr = a;
    }
    else if (dt.is(py::dtype::of<cfloat_t>()))
    {
        if (!(a.flags() & py::array::c_style))
            throw std::invalid_argument("ndarray must be row-major (C-contiguous)");
        s.arr = a;
    }
    else
    {
        s.arr = py::array_t<cfloat_t, py::array::c_style | py::array::forcecast>::ensure(a);
        if (!s.arr)
            throw py::type_error(std::string(what) + " cannot be converted to complex128");
    }
    s.ndim = std::size_t(s.arr.ndim());
    if (s.ndim >= 1)
        s.dim = std::size_t(s.arr.shape(0));
    if (s.ndim == 2)
        s.n_states = std::size_t(s.arr.shape(1));
    return s;
}

template <class T> std::complex<T> const *cptr(States const &s)
{
    return static_cast<std::complex<T> const *>(s.arr.data());
}

template <class T> py::array_t<std::complex<T>> zeros(std::vector<py::ssize_t> const &shape)
{
    py::array_t<std::complex<T>> a(shape);
    std::fill(a.mutable_data(), a.mutable_data() + a.size(), std::complex<T>(0));
    return a;
}

std::vector<py::ssize_t> shape_of(States const &s)
{
    if (s.ndim == 1)
        return {py::ssize_t(s.dim)};
    return {py::ssize_t(s.dim), py::ssize_t(s.n_states)};
}

void check_ndim(States const &s, char const *fn, std::size_t max_ndim = 2)
{
    if (s.ndim < 1 || s.ndim > max_ndim)
        throw std::invalid_argument(std::string(fn) + ": expected 1 or 2 dimensions, got " + std::to_string(s.ndim));
}

std::vector<cfloat_t> coeff_vector(py::handle obj)
{
    py::array_t<cfloat_t, py::array::c_style | py::array::forcecast> a =
        py::array_t<cfloat_t, py::array::c_style | py::array::forcecast>::ensure(obj);
    if (!a)
        throw py::type_error("coefficients must be array-like of complex numbers");
    if (a.ndim() != 1)
        throw std::invalid_argument("coefficients must be one-dimensional");
    return std::vector<cfloat_t>(a.data(), a.data() + a.size());
}

std::vector<std::string> strings_of(std::vector<fp::PauliString> const &ps)
{
    std::vector<std::string> out;
    out.reserve(ps.size());
    for (auto const &p : ps)
        out.push_back(p.str());
    return out;
}

// ---- PauliString -----------------------------------------------------------

py::array ps_apply(fp::PauliString const &self, py::handle states_obj, cfloat_t c)
{
    States const s = as_states(states_obj);
    check_ndim(s, "apply");
    if (s.dim != self.dim())
        throw std::invalid_argument("[PauliString] states shape (" + std::to_string(s.dim) +
                                    ") must match the dimension of the operators (" + std::to_string(self.dim()) + ")");
    if (s.c64)
    {
        auto out = zeros<float>(shape_of(s));
        fp::detail::string_apply_batch<float>(self.device_ops(), 0, cptr<float>(s), out.mutable_data(), s.dim,
                                              s.n_states, std::complex<float>(c), false);
        return out;
    }
    auto out = zeros<double>(shape_of(s));
    fp::detail::string_apply_batch<double>(self.device_ops(), 0, cptr<double>(s), out.mutable_data(), s.dim, s.n_states,
                                           c, false);
    return out;
}

py::array ps_expectation(fp::PauliString const &self, py::handle states_obj, cfloat_t c)
{
    States const s = as_states(states_obj);
    check_ndim(s, "expectation_value");
    if (s.dim != self.dim())
        throw std::invalid_argument("[PauliString] states shape (" + std::to_string(s.dim) +
                                    ") must match the dimension of the operators (" + std::to_string(self.dim()) + ")");
    if (s.c64)
    {
        auto out = zeros<float>({py::ssize_t(s.n_states)});
        fp::detail::string_expectation<float>(self.device_ops(), 0, cptr<float>(s), out.mutable_data(), s.dim,
                                              s.n_states, std::complex<float>(c), false);
        return out;
    }
    auto out = zeros<double>({py::ssize_t(s.n_states)});
    fp::detail::string_expectation<double>(self.device_ops(), 0, cptr<double>(s), out.mutable_data(), s.dim,
                                           s.n_states, c, false);
    return out;
}

py::array dense_string(fp::PauliString const &self)
{
    std::size_t const d = self.dim();
    auto out = zeros<double>({py::ssize_t(d), py::ssize_t(d)});
    if (d)
        self.to_tensor<double>(fp::mdspan(out.mutable_data(), d, d));
    return out;
}

// ---- PauliOp ---------------------------------------------------------------

using Op = fp::PauliOp<float_type>;

py::array op_apply(Op const &self, py::handle states_obj)
{
    States const s = as_states(states_obj);
    check_ndim(s, "apply");
    if (s.dim != self.dim())
        throw std::invalid_argument("[PauliOp] state size must match the dimension of the operators");
    if (s.c64)
    {
        auto out = zeros<float>(shape_of(s));
        fp::detail::op_apply<float>(self.device_ops(), cptr<float>(s), out.mutable_data(), s.dim, s.n_states, false);
        return out;
    }
    auto out = zeros<double>(shape_of(s));
    fp::detail::op_apply<double>(self.device_ops(), cptr<double>(s), out.mutable_data(), s.dim, s.n_states, false);
    return out;
}

py::array op_expectation(Op const &self, py::handle states_obj)
{
    States const s = as_states(states_obj);
    check_ndim(s, "expectation_value");
    if (s.dim != self.dim())
        throw std::invalid_argument("[PauliOp] state size must match the dimension of the operators");
    if (s.c64)
    {
        auto out = zeros<float>({py::ssize_t(s.n_states)});
        fp::detail::op_expectation<float>(self.device_ops(), cptr<float>(s), out.mutable_data(), s.dim, s.n_states,
                                          false);
        return out;
    }
    auto out = zeros<double>({py::ssize_t(s.n_states)});
    fp::detail::op_expectation<double>(self.device_ops(), cptr<double>(s), out.mutable_data(), s.dim, s.n_states,
                                       false);
    return out;
}

py::array dense_op(Op const &self)
{
    std::size_t const d = self.dim();
    auto out = zeros<double>({py::ssize_t(d), py::ssize_t(d)});
    if (d)
        self.to_tensor(fp::mdspan(out.mutable_data(), d, d));
    return out;
}

// ---- SummedPauliOp ---------------------------------------------------------

using SOp = fp::SummedPauliOp<float_type>;
using SOp32 = fp::SummedPauliOp<float>;

SOp make_summed(std::vector<fp::PauliString> const &strings, py::handle coeffs_obj)
{
    auto a = py::array_t<cfloat_t, py::array::c_style | py::array::forcecast>::ensure(coeffs_obj);
    if (!a)
        throw py::type_error("coeffs must be array-like");
    if (a.ndim() != 2)
        throw std::invalid_argument("coeffs must be a 2-D (n_pauli_strings, n_operators) array");
    return SOp(strings, fp::mdspan(a.data(), std::size_t(a.shape(0)), std::size_t(a.shape(1))));
}

SOp32 to_float(SOp const &self)
{
    std::vector<std::complex<float>> raw(self.coeffs_raw.begin(), self.coeffs_raw.end());
    SOp32 o(self.pauli_strings, raw);
    o.contraction = self.contraction;
    return o;
}

py::array sop_expectation(SOp const &self, py::handle states_obj)
{
    States const s = as_states(states_obj);
    check_ndim(s, "expectation_value");
    if (s.dim != self.dim())
        throw std::invalid_argument("states must have the same dimension (" + std::to_string(s.dim) +
                                    ") as the SummedPauliOp (" + std::to_string(self.dim()) + ")");
    std::vector<py::ssize_t> shape = s.ndim == 1 ? std::vector<py::ssize_t>{py::ssize_t(self.n_operators())}
                                                 : std::vector<py::ssize_t>{py::ssize_t(self.n_operators()),
                                                                            py::ssize_t(s.n_states)};
    if (s.c64)
    {
        auto out = zeros<float>(shape);
        to_float(self).expectation_value_raw(cptr<float>(s), out.mutable_data(), s.n_states, false);
        return out;
    }
    auto out = zeros<double>(shape);
    self.expectation_value_raw(cptr<double>(s), out.mutable_data(), s.n_states, false);
    return out;
}

py::array sop_apply(SOp const &self, py::handle states_obj)
{
    States const s = as_states(states_obj);
    check_ndim(s, "apply");
    if (s.dim != self.dim())
        throw std::invalid_argument("state size must match the dimension of the operators");
    if (s.c64)
    {
        auto out = zeros<float>(shape_of(s));
        to_float(self).apply_raw(cptr<float>(s), out.mutable_data(), s.n_states, false);
        return out;
    }
    auto out = zeros<double>(shape_of(s));
    self.apply_raw(cptr<double>(s), out.mutable_data(), s.n_states, false);
    return out;
}

py::array sop_apply_weighted(SOp const &self, py::handle states_obj, py::handle data_obj)
{
    States const s = as_states(states_obj);
    check_ndim(s, "apply_weighted");
    if (s.dim != self.dim())
        throw std::invalid_argument("state size must match the dimension of the operators");
    if (s.c64)
    {
        auto d = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(data_obj);
        if (!d)
            throw py::type_error("data must be a real array");
        std::size_t const rows = d.ndim() >= 1 ? std::size_t(d.shape(0)) : 0;
        std::size_t const cols = d.ndim() == 2 ? std::size_t(d.shape(1)) : 1;
        if (rows != self.n_operators() || cols != s.n_states || std::size_t(d.ndim()) != s.ndim)
            throw std::invalid_argument("data(k,t) must have the same number of operators as the SummedPauliOp and the "
                                        "same number of states as the input states");
        auto out = zeros<float>(shape_of(s));
        to_float(self).apply_weighted_raw(cptr<float>(s), d.data(), out.mutable_data(), s.n_states, false);
        return out;
    }
    auto d = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(data_obj);
    if (!d)
        throw py::type_error("data must be a real array");
    std::size_t const rows = d.ndim() >= 1 ? std::size_t(d.shape(0)) : 0;
    std::size_t const cols = d.ndim() == 2 ? std::size_t(d.shape(1)) : 1;
    if (rows != self.n_operators() || cols != s.n_states || std::size_t(d.ndim()) != s.ndim)
        throw std::invalid_argument("data(k,t) must have the same number of operators as the SummedPauliOp and the "
                                    "same number of states as the input states");
    auto out = zeros<double>(shape_of(s));
    self.apply_weighted_raw(cptr<double>(s), d.data(), out.mutable_data(), s.n_states, false);
    return out;
}

py::array sop_to_tensor(SOp const &self)
{
    std::size_t const d = self.dim(), K = self.n_operators();
    auto out = zeros<double>({py::ssize_t(K), py::ssize_t(d), py::ssize_t(d)});
    self.to_tensor(fp::mdspan(out.mutable_data(), K, d, d));
    return out;
}

py::array sop_coeffs(SOp const &self)
{
    std::size_t const N = self.coeffs.extent(0), K = self.coeffs.extent(1);
    auto out = zeros<double>({py::ssize_t(K), py::ssize_t(N)});
    auto *o = out.mutable_data();
    for (std::size_t i = 0; i < N; ++i)
        for (std::size_t j = 0; j < K; ++j)
            o[j * N + i] = self.coeffs(i, j);
    return out;
}

void sop_set_coeffs(SOp &self, py::handle obj)
{
    auto a = py::array_t<cfloat_t, py::array::c_style | py::array::forcecast>::ensure(obj);
    if (!a || a.ndim() != 2 || std::size_t(a.shape(0)) != self.n_operators() ||
        std::size_t(a.shape(1)) != self.n_pauli_strings())
        throw std::invalid_argument("The shape of provided coeffs must match the number of operators and PauliStrings");
    std::size_t const N = self.coeffs.extent(0), K = self.coeffs.extent(1);
    for (std::size_t i = 0; i < N; ++i)
        for (std::size_t j = 0; j < K; ++j)
            self.coeffs(i, j) = a.data()[j * N + i];
}

} // namespace

PYBIND11_MODULE(_fast_pauli, m)
{
    m.doc() = "fast_pauli on MI355X (gfx950): Pauli strings and operators applied on the GPU";
    py::register_exception_translator([](std::exception_ptr p) {
        try
        {
            if (p)
                std::rethrow_exception(p);
        }
        catch (std::invalid_argument const &e)
        {
            PyErr_SetString(PyExc_ValueError, e.what());
        }
    });

    // ---------------------------------------------------------------- Pauli
    py::class_<fp::Pauli>(m, "Pauli",
                          "A class for efficient representation of a 2x2 Pauli Matrix sigma_i in {I, X, Y, Z}")
        .def(py::init<>(), "Default constructor to initialize with identity matrix.")
        .def(py::init([](int code) { return fp::Pauli(code); }), "code"_a,
             "Constructor given a numeric code (0: I, 1: X, 2: Y, 3: Z).")
        .def(py::init([](std::string const &symbol) {
                 if (symbol.size() != 1)
                     throw py::type_error("Pauli symbol must be a single character");
                 return fp::Pauli(symbol[0]);
             }),
             "symbol"_a, "Constructor given Pauli matrix symbol (I, X, Y or Z).")
        .def(
            "__matmul__",
            [](fp::Pauli const &a, fp::Pauli const &b) {
                auto [ph, p] = a * b;
                return py::make_tuple(ph, p);
            },
            py::is_operator(), "Returns matrix product of two Paulis as a tuple of phase and new Pauli object.")
        .def(
            "to_tensor",
            [](fp::Pauli const &self) {
                auto out = zeros<double>({2, 2});
                self.to_tensor<double>(fp::mdspan(out.mutable_data(), 2, 2));
                return out;
            },
            "Returns a dense representation of Pauli object as a 2x2 matrix.")
        .def("clone", [](fp::Pauli const &self) { return fp::Pauli(self); }, "Returns a copy of the Pauli object.")
        .def("__str__", [](fp::Pauli const &self) { return fp::to_string(self); })
        .def("__repr__", [](fp::Pauli const &self) { return "Pauli('" + fp::to_string(self) + "')"; })
        .def(py::pickle([](fp::Pauli const &self) { return int(self.code); },
                        [](int code) { return fp::Pauli(code); }));

    // ----------------------------------------------------------- PauliString
    py::class_<fp::PauliString>(m, "PauliString",
                                "A Pauli string: a tensor product of Pauli matrices sigma_i in {I, X, Y, Z}")
        .def(py::init<>(), "Default constructor to initialize with empty string.")
        .def(py::init<std::string const &>(), "string"_a,
             "Constructs a PauliString from a string of I, X, Y, Z characters.")
        .def(py::init<std::vector<fp::Pauli>>(), "paulis"_a,
             "Constructs a PauliString from a list of Pauli objects.")
        .def("__str__", [](fp::PauliString const &self) { return self.str(); })
        .def("__repr__", [](fp::PauliString const &self) { return "PauliString('" + self.str() + "')"; })
        .def(
            "__matmul__",
            [](fp::PauliString const &self, fp::PauliString const &rhs) {
                auto [ph, p] = self * rhs;
                return py::make_tuple(ph, p);
            },
            py::is_operator(), "Returns matrix product of two pauli strings and their phase as a pair.")
        .def(
            "__add__", [](fp::PauliString const &self, fp::PauliString const &other) { return Op({self, other}); },
            py::is_operator(), "Returns the sum of two Pauli strings in a form of PauliOp object.")
        .def(
            "__sub__",
            [](fp::PauliString const &self, fp::PauliString const &other) {
                return Op(std::vector<cfloat_t>{1.0, -1.0}, {self, other});
            },
            py::is_operator(), "Returns the difference of two Pauli strings in a form of PauliOp object.")
        .def_property_readonly("n_qubits", &fp::PauliString::n_qubits,
                               "int: The number of qubits in PauliString (i.e. number of Pauli Matrices in tensor "
                               "product)")
        .def_property_readonly("dim", &fp::PauliString::dim, "int: The dimension of PauliString 2^n")
        .def_property_readonly(
            "weight", [](fp::PauliString const &self) { return int(self.weight); },
            "int: The weight of PauliString (i.e. number of non-identity Pauli matrices in it)")
        .def("apply", &ps_apply, "states"_a, "coeff"_a = cfloat_t{1.0},
             "Apply c * P to a 1-D (n_dims,) state or a 2-D (n_dims, n_states) batch of states stored as columns.")
        .def("expectation_value", &ps_expectation, "states"_a, "coeff"_a = cfloat_t{1.0},
             "c * <psi_t|P|psi_t> for a 1-D state or each column of a 2-D batch; returns (n_states,).")
        .def("to_tensor", &dense_string, "Returns a dense representation of PauliString.")
        .def("clone", [](fp::PauliString const &self) { return fp::PauliString(self); },
             "Returns a copy of the PauliString object.")
        .def(py::pickle([](fp::PauliString const &self) { return self.paulis; },
                        [](std::vector<fp::Pauli> paulis) { return fp::PauliString(std::move(paulis)); }));

    // --------------------------------------------------------------- PauliOp
    py::class_<Op>(m, "PauliOp", "A Pauli operator A = sum_j h_j P_j")
        .def(py::init<>())
        .def(py::init<std::vector<std::string> const &>(), "pauli_strings"_a)
        .def(py::init<std::vector<fp::PauliString>>(), "pauli_strings"_a)
        .def(py::init([](py::array coefficients, std::vector<fp::PauliString> const &pauli_strings) {
                 return Op(coeff_vector(coefficients), pauli_strings);
             }),
             "coefficients"_a, "pauli_strings"_a)
        .def(py::init([](std::vector<cfloat_t> coefficients, std::vector<fp::PauliString> const &pauli_strings) {
                 return Op(std::move(coefficients), pauli_strings);
             }),
             "coefficients"_a, "pauli_strings"_a)
        .def(py::init([](py::object coefficients, std::vector<std::string> const &strings) {
                 std::vector<fp::PauliString> ps;
                 ps.reserve(strings.size());
                 for (auto const &s : strings)
                     ps.emplace_back(s);
                 return Op(coeff_vector(coefficients), ps);
             }),
             "coefficients"_a, "pauli_strings"_a)
        .def(
            "__matmul__", [](Op const &self, Op const &rhs) { return self * rhs; }, py::is_operator())
        .def(
            "__matmul__", [](Op const &self, fp::PauliString const &rhs) { return self * rhs; }, py::is_operator())
        .def(
            "__rmatmul__", [](Op const &self, fp::PauliString const &lhs) { return lhs * self; }, py::is_operator())
        .def(
            "__mul__",
            [](Op const &self, cfloat_t rhs) {
                Op r(self);
                r.scale(rhs);
                return r;
            },
            py::is_operator())
        .def(
            "__rmul__",
            [](Op const &self, cfloat_t lhs) {
                Op r(self);
                r.scale(lhs);
                return r;
            },
            py::is_operator())
        .def(
            "__imul__",
            [](Op &self, cfloat_t f) -> Op & {
                self.scale(f);
                return self;
            },
            py::is_operator())
        .def(
            "__add__",
            [](Op const &l, Op const &r) {
                Op o(l);
                o.extend(r);
                return o;
            },
            py::is_operator())
        .def(
            "__add__",
            [](Op const &l, fp::PauliString const &r) {
                Op o(l);
                o.extend(r, 1);
                return o;
            },
            py::is_operator())
        .def(
            "__radd__",
            [](Op const &self, fp::PauliString const &l) {
                Op o(self);
                o.extend(l, 1);
                return o;
            },
            py::is_operator())
        .def(
            "__iadd__",
            [](Op &self, Op const &other) -> Op & {
                self.extend(other);
                return self;
            },
            py::is_operator())
        .def(
            "__iadd__",
            [](Op &self, fp::PauliString const &other) -> Op & {
                self.extend(other, 1);
                return self;
            },
            py::is_operator())
        .def(
            "__sub__",
            [](Op const &l, Op const &r) {
                Op o(l);
                o.extend(-r);
                return o;
            },
            py::is_operator())
        .def(
            "__sub__",
            [](Op const &l, fp::PauliString const &r) {
                Op o(l);
                o.extend(r, -1);
                return o;
            },
            py::is_operator())
        .def(
            "__rsub__",
            [](Op const &self, fp::PauliString const &l) {
                Op o(-self);
                o.extend(l, 1);
                return o;
            },
            py::is_operator())
        .def(
            "__isub__",
            [](Op &self, Op const &other) -> Op & {
                self.extend(-other);
                return self;
            },
            py::is_operator())
        .def(
            "__isub__",
            [](Op &self, fp::PauliString const &other) -> Op & {
                self.extend(other, -1);
                return self;
            },
            py::is_operator())
        .def(
            "extend", [](Op &self, Op const &other) { self.extend(other); }, "other"_a)
        .def(
            "extend",
            [](Op &self, fp::PauliString const &other, cfloat_t multiplier, bool dedupe) {
                self.extend(other, multiplier, dedupe);
            },
            "other"_a, "multiplier"_a, "dedupe"_a = true)
        .def_property_readonly("dim", &Op::dim)
        .def_property_readonly("n_qubits", &Op::n_qubits)
        .def_property_readonly("n_pauli_strings", &Op::n_pauli_strings)
        .def_property_readonly("coeffs", [](Op const &self) { return self.coeffs; })
        .def_property_readonly("pauli_strings", [](Op const &self) { return self.pauli_strings; })
        .def_property_readonly("pauli_strings_as_str", [](Op const &self) { return strings_of(self.pauli_strings); })
        .def(
            "scale", [](Op &self, cfloat_t factor) { self.scale(factor); }, "factor"_a)
        .def(
            "scale",
            [](Op &self, py::array factors) {
                auto v = coeff_vector(factors);
                self.scale(fp::mdspan(v.data(), v.size()));
            },
            "factors"_a)
        .def("apply", &op_apply, "states"_a,
             "(sum_j h_j P_j) applied to a 1-D state or to each column of a 2-D (n_dims, n_states) batch.")
        .def("expectation_value", &op_expectation, "states"_a,
             "<psi_t| sum_j h_j P_j |psi_t> for a 1-D state or each column of a 2-D batch; returns (n_states,).")
        .def("to_tensor", &dense_op)
        .def("clone", [](Op const &self) { return Op(self); })
        .def(py::pickle([](Op const &self) { return py::make_tuple(self.coeffs, self.pauli_strings); },
                        [](py::tuple t) {
                            return Op(t[0].cast<std::vector<cfloat_t>>(), t[1].cast<std::vector<fp::PauliString>>());
                        }));

    // --------------------------------------------------------- SummedPauliOp
    py::class_<SOp>(m, "SummedPauliOp", "K Pauli operators A_k = sum_i h_ik P_i sharing one set of Pauli strings")
        .def(py::init<>())
        .def(py::init([](std::vector<fp::PauliString> const &pauli_strings, py::handle coeffs) {
                 return make_summed(pauli_strings, coeffs);
             }),
             "pauli_strings"_a, "coeffs"_a)
        .def(py::init([](std::vector<std::string> const &pauli_strings, py::handle coeffs) {
                 std::vector<fp::PauliString> ps;
                 for (auto const &s : pauli_strings)
                     ps.emplace_back(s);
                 return make_summed(ps, coeffs);
             }),
             "pauli_strings"_a, "coeffs"_a)
        .def_property_readonly("dim", &SOp::dim)
        .def_property_readonly("n_operators", &SOp::n_operators)
        .def_property_readonly("n_pauli_strings", &SOp::n_pauli_strings)
        .def_property("coeffs", &sop_coeffs, &sop_set_coeffs,
                      "np.ndarray: coefficients with shape (n_operators, n_pauli_strings)")
        .def_property_readonly("pauli_strings", [](SOp const &self) { return self.pauli_strings; })
        .def_property_readonly("pauli_strings_as_str", [](SOp const &self) { return strings_of(self.pauli_strings); })
        .def_property(
            "contraction", [](SOp const &self) { return self.contraction; },
            [](SOp &self, int mode) {
                if (mode < FP_CONTRACT_FP64 || mode > FP_CONTRACT_AUTO)
                    throw std::invalid_argument("unknown contraction mode");
                self.contraction = mode;
            },
            "int: precision of the expectation-value contraction (0 FP64, 1 FP32, 2 BF16x3, 3 auto)")
        .def("apply", &sop_apply, "states"_a)
        .def("apply_weighted", &sop_apply_weighted, "states"_a, "data"_a)
        .def("expectation_value", &sop_expectation, "states"_a)
        .def("to_tensor", &sop_to_tensor)
        .def("clone", [](SOp const &self) { return SOp(self.pauli_strings, self.coeffs_raw); })
        .def("split", &SOp::split)
        .def("square", &SOp::square)
        .def(py::pickle([](SOp const &self) { return py::make_tuple(self.pauli_strings, self.coeffs_raw); },
                        [](py::tuple t) {
                            return SOp(t[0].cast<std::vector<fp::PauliString>>(), t[1].cast<std::vector<cfloat_t>>());
                        }));

    // --------------------------------------------------------------- helpers
    auto h = m.def_submodule("helpers");
    h.def("get_nontrivial_paulis", &fp::get_nontrivial_paulis, "weight"_a);
    h.def("calculate_pauli_strings", &fp::calculate_pauli_strings, "n_qubits"_a, "weight"_a);
    h.def("calculate_pauli_strings_max_weight", &fp::calculate_pauli_strings_max_weight, "n_qubits"_a, "weight"_a);
    h.def(
        "pauli_string_sparse_repr",
        [](std::vector<fp::Pauli> const &paulis) {
            auto [k, v] = fp::get_sparse_repr<float_type>(paulis);
            return py::make_tuple(k, v);
        },
        "paulis"_a);

    m.attr("FP_CONTRACT_FP64") = int(FP_CONTRACT_FP64);
    m.attr("FP_CONTRACT_FP32") = int(FP_CONTRACT_FP32);
    m.attr("FP_CONTRACT_BF16X3") = int(FP_CONTRACT_BF16X3);
    m.attr("FP_CONTRACT_AUTO") = int(FP_CONTRACT_AUTO);
}
