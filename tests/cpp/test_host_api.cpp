// C++ host API checks, run on the GPU: the reference's C++ call forms
// (fast_pauli/cpp/tests/test_pauli_string.cpp, test_pauli_op.cpp,
// test_summed_pauli_op.cpp) against dense matrix products computed here.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <execution>
#include <random>
#include <vector>

#include "fast_pauli/fast_pauli.hpp"

using namespace fast_pauli;
using cd = std::complex<double>;

static int failures = 0;
#define CHECK(cond)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            std::printf("CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__);                                      \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

static std::vector<cd> dense(PauliString const &ps)
{
    std::size_t const d = ps.dim();
    std::vector<cd> m(d * d);
    ps.to_tensor<double>(mdspan(m.data(), d, d));
    return m;
}

static std::vector<cd> rand_vec(std::size_t n, unsigned seed)
{
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> u(0, 1);
    std::vector<cd> v(n);
    for (auto &x : v)
        x = cd(u(g), u(g));
    return v;
}

int main()
{
    std::vector<std::string> const strings = {"IXYZ", "YYIX", "XXYIYZ", "IZIXYYZ", "IZIXYYZIXYZ"};
    for (auto const &s : strings)
    {
        PauliString ps(s);
        std::size_t const d = ps.dim(), S = 5;
        auto const m = dense(ps);
        auto const psi = rand_vec(d * S, 18);
        // apply_batch accumulates into new_states (seq and par policies)
        std::vector<cd> out(d * S, cd(1.0, 0.0)), out_par(d * S, cd(1.0, 0.0));
        cd const c(0.5, -1.5);
        ps.apply_batch(mdspan(out.data(), d, S), mdspan(psi.data(), d, S), c);
        ps.apply_batch(std::execution::par, mdspan(out_par.data(), d, S), mdspan(psi.data(), d, S), c);
        double err = 0;
        for (std::size_t i = 0; i < d; ++i)
            for (std::size_t t = 0; t < S; ++t)
            {
                cd acc = 1.0;
                for (std::size_t j = 0; j < d; ++j)
                    acc += c * m[i * d + j] * psi[j * S + t];
                err = std::max(err, std::abs(acc - out[i * S + t]));
                CHECK(out[i * S + t] == out_par[i * S + t]);
            }
        CHECK(err < 1e-12);
        // 1-D apply
        std::vector<cd> o1(d), p1(psi.begin(), psi.begin() + d);
        ps.apply(mdspan(o1.data(), d), mdspan(p1.data(), d));
        for (std::size_t i = 0; i < d; ++i)
        {
            cd acc = 0;
            for (std::size_t j = 0; j < d; ++j)
                acc += m[i * d + j] * p1[j];
            CHECK(std::abs(acc - o1[i]) < 1e-12);
        }
        // expectation values accumulate too
        std::vector<cd> ev(S, cd(2.0, 0.0));
        ps.expectation_value(std::execution::par, mdspan(ev.data(), S), mdspan(psi.data(), d, S));
        for (std::size_t t = 0; t < S; ++t)
        {
            cd acc = 2.0;
            for (std::size_t i = 0; i < d; ++i)
                for (std::size_t j = 0; j < d; ++j)
                    acc += std::conj(psi[i * S + t]) * m[i * d + j] * psi[j * S + t];
            CHECK(std::abs(acc - ev[t]) < 1e-11);
        }
    }
    // PauliOp: bad shapes throw std::invalid_argument (test_pauli_op.cpp:101-117)
    {
        PauliOp<double> op(std::vector<std::string>{"XYZ", "ZZI"});
        std::vector<cd> a(16), b(16);
        bool threw = false;
        try
        {
            op.apply(mdspan(a.data(), 16, 1), mdspan(b.data(), 16, 1));
        }
        catch (std::invalid_argument const &)
        {
            threw = true;
        }
        CHECK(threw);
    }
    // PauliOp apply == sum of PauliString applies, bit-exact (test_pauli_op.cpp:159-200)
    {
        std::vector<cd> const state = {{0.6514, 0.8887}, {0.0903, 0.8848}, {0.8130, 0.6729}, {0.9854, 0.9497},
                                       {0.0410, 0.5132}, {0.7920, 0.6519}, {0.4683, 0.5708}, {0.2148, 0.4561},
                                       {0.5400, 0.9932}, {0.9736, 0.0449}, {0.2979, 0.9834}, {0.4346, 0.5635},
                                       {0.0767, 0.3516}, {0.4873, 0.4097}, {0.1543, 0.7329}, {0.4121, 0.5059}};
        std::vector<PauliString> ps = {"IXYZ", "ZZXY", "YIIY"};
        std::vector<cd> coeffs = {1.0, {0.0, 2.0}, {-0.5, 0.25}};
        PauliOp<double> op(coeffs, ps);
        std::vector<cd> res(16), expected(16);
        op.apply(std::execution::par, mdspan(res.data(), 16), mdspan(state.data(), 16));
        for (std::size_t s = 0; s < ps.size(); ++s)
            ps[s].apply(mdspan(expected.data(), 16), mdspan(state.data(), 16), coeffs[s]);
        for (std::size_t i = 0; i < 16; ++i)
            CHECK(res[i] == expected[i]);
        // expectation of 16 copies of I on a state = 16 * <psi|psi> (test_pauli_op.cpp:228-267)
        PauliOp<double> ids(std::vector<std::string>(16, "IIII"));
        std::vector<cd> ev(1);
        ids.expectation_value(mdspan(ev.data(), 1), mdspan(state.data(), 16, 1));
        double nrm = 0;
        for (auto const &v : state)
            nrm += std::norm(v);
        CHECK(std::abs(ev[0] - 16.0 * nrm) < 1e-12);
    }
    // SummedPauliOp expectation == sum_j h(j,k) <P_j> (test_summed_pauli_op.cpp:311-363)
    {
        std::size_t const n = 6, S = 7;
        auto const strs = calculate_pauli_strings_max_weight(n, 2);
        std::size_t const N = strs.size(), K = 4, d = std::size_t(1) << n;
        auto const h = rand_vec(N * K, 3);
        SummedPauliOp<double> sop(strs, mdspan(h.data(), N, K));
        auto const psi = rand_vec(d * S, 4);
        std::vector<cd> out(K * S);
        sop.expectation_value(std::execution::par, mdspan(out.data(), K, S), mdspan(psi.data(), d, S));
        std::vector<cd> ev(N * S);
        for (std::size_t j = 0; j < N; ++j)
            strs[j].expectation_value(mdspan(ev.data() + j * S, S), mdspan(psi.data(), d, S));
        for (std::size_t k = 0; k < K; ++k)
            for (std::size_t t = 0; t < S; ++t)
            {
                cd acc = 0;
                for (std::size_t j = 0; j < N; ++j)
                    acc += h[j * K + k] * ev[j * S + t];
                CHECK(std::abs(acc - out[k * S + t]) < 1e-10 * std::max(1.0, std::abs(acc)));
            }
    }
    // float instantiation (complex64 fast path)
    {
        PauliString ps("XYZZY");
        std::size_t const d = ps.dim(), S = 4;
        std::vector<std::complex<float>> psi(d * S), out(d * S);
        for (std::size_t i = 0; i < psi.size(); ++i)
            psi[i] = {float(i % 7) * 0.25f, float(i % 5) * -0.5f};
        ps.apply_batch(mdspan(out.data(), d, S), mdspan(psi.data(), d, S), std::complex<float>(1.0f));
        auto const m = dense(ps);
        for (std::size_t i = 0; i < d; ++i)
            for (std::size_t t = 0; t < S; ++t)
            {
                std::complex<double> acc = 0;
                for (std::size_t j = 0; j < d; ++j)
                    acc += m[i * d + j] * std::complex<double>(psi[j * S + t]);
                CHECK(std::abs(acc - std::complex<double>(out[i * S + t])) < 1e-6);
            }
    }
    // to_tensor into device memory runs on the device and equals the host form element by element
    {
        PauliString ps("XYZIY");
        std::size_t const d = ps.dim();
        std::vector<cd> host(d * d, cd(0.25, -0.5)), dev(d * d);
        hip::Buffer buf(d * d * sizeof(cd));
        hip::h2d(buf.get(), host.data(), d * d * sizeof(cd));
        ps.to_tensor<double>(mdspan(host.data(), d, d));
        ps.to_tensor<double>(mdspan(static_cast<cd *>(buf.get()), d, d));
        hip::d2h(dev.data(), buf.get(), d * d * sizeof(cd));
        for (std::size_t i = 0; i < d * d; ++i)
            CHECK(dev[i] == host[i]);

        std::vector<PauliString> const strs = {PauliString("XYZIY"), PauliString("ZZIXX"), PauliString("YZIXX"),
                                               PauliString("IIIII")};
        std::vector<cd> const cf = {cd(0.5, -1.0), cd(2.0, 0.0), cd(0.0, -0.75), cd(-1.25, 0.5)};
        PauliOp<double> op(cf, strs);
        std::vector<cd> h2 = rand_vec(d * d, 7), d2(d * d);
        hip::Buffer b2(d * d * sizeof(cd));
        hip::h2d(b2.get(), h2.data(), d * d * sizeof(cd));
        op.to_tensor(mdspan(h2.data(), d, d));
        op.to_tensor(mdspan(static_cast<cd *>(b2.get()), d, d));
        hip::d2h(d2.data(), b2.get(), d * d * sizeof(cd));
        for (std::size_t i = 0; i < d * d; ++i)
            CHECK(d2[i] == h2[i]);

        std::size_t const K = 3;
        auto const hk = rand_vec(strs.size() * K, 9);
        SummedPauliOp<double> sop(strs, hk); // coefficients (n_strings x K), row-major
        std::vector<cd> h3(K * d * d, cd(0.0, 0.0)), d3(K * d * d);
        hip::Buffer b3(K * d * d * sizeof(cd));
        hip::h2d(b3.get(), h3.data(), K * d * d * sizeof(cd));
        sop.to_tensor(mdspan(h3.data(), K, d, d));
        sop.to_tensor(mdspan(static_cast<cd *>(b3.get()), K, d, d));
        hip::d2h(d3.data(), b3.get(), K * d * d * sizeof(cd));
        for (std::size_t i = 0; i < K * d * d; ++i)
            CHECK(d3[i] == h3[i]);
    }
    std::printf("host API checks: %s (%d failures)\n", failures ? "FAILED" : "passed", failures);
    return failures ? 1 : 0;
}
