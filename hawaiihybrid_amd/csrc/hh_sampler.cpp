// Host-side initial-condition sampler and Mersenne Twister of the engine.
//
// The initial conditions are a few transcendental functions of a sequential Mersenne
// Twister stream.  They run on the host on purpose: the reference evaluates them with the
// host libm (log, cos, sin, acos, pow, exp), and reproducing those last-ulp results is
// what makes the sampled phase points bit-identical to the reference's from the same seed.
// The cost is ~15 draws and ~10 libm calls per trajectory (milliseconds per 1e5
// trajectories), and the result is uploaded once per job.
//
// Compiled without floating-point contraction (the reference's default build, FLAGS_DEBUG
// -O0 in its Makefile, performs none).
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/hawaii_gpu.h"
#include "ch4co2_standin.h"
#include "co2ar_tables.h"

namespace {

// ---------------------------------------------------------------- MT19937 with mtwist seeding
// mt_seed32 (thirdparty/mtwist.c:372-406): state from the 69069 LCG, seed 0 -> 4357, refresh
// at once; mt_lrand / mt_drand (thirdparty/mtwist.h:560-634): tempered words in generation
// order, drand = word * 2^-32.
class Mt {
  public:
    explicit Mt(uint32_t seed)
    {
        if (seed == 0) seed = 4357u;
        s_[0] = seed;
        for (int i = 1; i < kN; ++i) s_[i] = 69069u * s_[i - 1];
        twist();
    }
    uint32_t next()
    {
        if (pos_ == kN) twist();
        uint32_t y = s_[pos_++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        return y ^ (y >> 18);
    }
    double drand()
    {
        ++draws;
        return static_cast<double>(next()) * (1.0 / 4294967296.0);
    }
    long draws = 0;

  private:
    static constexpr int kN = 624, kM = 397;
    void twist()
    {
        int i = 0;
        for (; i < kN - kM; ++i) s_[i] = s_[i + kM] ^ mix(s_[i], s_[i + 1]);
        for (; i < kN - 1; ++i) s_[i] = s_[i + kM - kN] ^ mix(s_[i], s_[i + 1]);
        s_[kN - 1] = s_[kM - 1] ^ mix(s_[kN - 1], s_[0]);
        pos_ = 0;
    }
    static uint32_t mix(uint32_t a, uint32_t b)
    {
        const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
        return (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    uint32_t s_[kN];
    int pos_ = kN;
};

constexpr double kHkT = 4.3597447222071e-18 / 1.380649e-23;   // Hartree / Boltzmann (src/constants.h:24)
constexpr double kHTOCM = 2.1947463136320e5;

// ---------------------------------------------------------------- built-in host surfaces
// Used by the rejection step when the caller supplies no plugin pes_lab.
double hear_V(double R)
{
    const double A = 0.2303058634e2, a1 = 0.163051044e1, a2 = 0.4680016e-1, b = 0.16274792500000e1;
    const double C[6] = {0.940835513e1, 0.165523018e3, 0.379715796e4, 0.1165179999e6, 0.466258000e7, 0.236861000e9};
    const double x = b * R, e = std::exp(-x);
    double term = 1.0, sum = 1.0, V = A * std::exp(-a1 * R - a2 * R * R), R2n = R * R * R * R;
    int k = 0;
    for (int n = 3; n <= 8; ++n) {
        R2n *= R * R;
        while (k < 2 * n) {
            ++k;
            term *= x / k;
            sum += term;
        }
        V -= C[n - 3] * (1.0 - e * sum) / R2n;
    }
    return V;
}

double co2ar_V(const double* q)
{
    // q = [Phi, Theta, R, phi1, theta1]; cos(theta_m) = Rhat . u
    const double c0 = std::sin(q[1]) * std::cos(q[0]) * std::sin(q[4]) * std::cos(q[3]) +
                      std::sin(q[1]) * std::sin(q[0]) * std::sin(q[4]) * std::sin(q[3]) +
                      std::cos(q[1]) * std::cos(q[4]);
    const double c = c0 > 1.0 ? 1.0 : (c0 < -1.0 ? -1.0 : c0);
    double P[11];
    P[0] = 1.0;
    P[1] = c;
    for (int l = 1; l < 10; ++l) P[l + 1] = ((2 * l + 1) * c * P[l] - l * P[l - 1]) / (l + 1);
    const double R = q[2];
    int b = 0;
    while (b < HH_CO2AR_NBRANCH - 1 && !(R < HH_CO2AR_BOUNDS[b])) ++b;
    double V = 0.0;
    if (b == 0) {
        for (int s = 0; s < 6; ++s) V += HH_CO2AR_PES_EXP[s][0] * std::exp(HH_CO2AR_PES_EXP[s][1] * R) * P[2 * s];
    } else {
        const double x = R - HH_CO2AR_CENTER[b], si = 1.0 / R, s4 = si * si * si * si;
        for (int s = 0; s < 6; ++s) {
            const double* cf = HH_CO2AR_PES[b][s];
            V += (((cf[3] * x + cf[2]) * x + cf[1]) * x + cf[0] + s4 * (((cf[7] * si + cf[6]) * si + cf[5]) * si + cf[4])) *
                 P[2 * s];
        }
    }
    return V;
}

// CH4-CO2 stand-in surface (ch4co2_standin.h) at lab q[8]: the directions of the intermolecular
// vector and the CO2 axis in the frame of CH4 (S1 = Sz(psi1) Sx(theta1) Sz(phi1) applied to the
// reference's lab vectors, src/angles_handler.cpp:357-419), then the site sum
double ch4co2_V(const double* q)
{
    using namespace hh::ch4co2;
    const double v[2][3] = {{-std::sin(q[1]) * std::cos(q[0]), -std::sin(q[1]) * std::sin(q[0]), std::cos(q[1])},
                            {-std::sin(q[7]) * std::cos(q[6]), -std::sin(q[7]) * std::sin(q[6]), std::cos(q[7])}};
    const double sf = std::sin(q[3]), cf = std::cos(q[3]), s1 = std::sin(q[4]), c1 = std::cos(q[4]);
    const double sp = std::sin(q[5]), cp = std::cos(q[5]);
    double ab[2][3];
    for (int k = 0; k < 2; ++k) {
        const double w0 = cf * v[k][0] + sf * v[k][1], w1 = -sf * v[k][0] + cf * v[k][1], w2 = v[k][2];
        const double x1 = c1 * w1 + s1 * w2, x2 = -s1 * w1 + c1 * w2;
        ab[k][0] = cp * w0 + sp * x1;
        ab[k][1] = -sp * w0 + cp * x1;
        ab[k][2] = x2;
    }
    const double R = q[2], hs = kRCH / std::sqrt(3.0);
    double V = 0.0;
    for (int j = 0; j < 3; ++j) {
        const double sg = j == 0 ? 0.0 : (j == 1 ? 1.0 : -1.0);
        const int tj = j == 0 ? 2 : 3;
        for (int i = 0; i < 5; ++i) {
            const int ti = i == 0 ? 0 : 1;
            const double sig = 0.5 * (kSigma[ti] + kSigma[tj]) * kAng;
            const double e4 = 4.0 * kWellScale * kKelvinHartree * std::sqrt(kEpsilon[ti] * kEpsilon[tj]);
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double h = i == 0 ? 0.0 : hs * kH[i - 1][c];
                const double d = R * ab[0][c] + sg * kRCO * ab[1][c] - h;
                r2 += d * d;
            }
            const double sr2 = sig * sig / r2, sr6 = sr2 * sr2 * sr2;
            V += e4 * (sr6 * sr6 - sr6);
        }
    }
    return V;
}

struct Sys {
    int id;
    double mu, I;
    const double* IR;
    hh_pes_fn pes;
    bool linear() const { return id == HH_SYS_CO2AR || id == HH_SYS_N2AR; }   // LINEAR_MOLECULE + ATOM
    bool rotor_linear() const { return id == HH_SYS_CH4CO2; }                 // ROTOR + LINEAR_MOLECULE
    int qp() const { return rotor_linear() ? 16 : linear() ? 10 : 6; }
    double V(const double* y) const
    {
        double q[8];
        const int nq = qp() / 2;
        for (int i = 0; i < nq; ++i) q[i] = y[2 * i];
        if (pes) return pes(q);
        if (rotor_linear()) return ch4co2_V(q);
        return id == HH_SYS_CO2AR ? co2ar_V(q) : hear_V(q[2]);
    }
    // kinetic_energy, src/hawaii.c:1041-1137 (same operation order)
    double K(const double* y) const
    {
        const double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5];
        const double sinTheta = std::sin(Theta);
        double k1 = pR * pR / (2.0 * mu) + pTheta * pTheta / (2.0 * mu * R * R) +
                    pPhi * pPhi / (2.0 * mu * R * R * sinTheta * sinTheta);
        if (linear()) {
            const double st = std::sin(y[8]);
            k1 = k1 + (y[9] * y[9] / (2.0 * I) + y[7] * y[7] / (2.0 * I * st * st));
        }
        if (rotor_linear()) {
            // KIN1 + KIN2 (ROTOR monomer 1) + KIN3 (LINEAR_MOLECULE monomer 2)
            const double pphi1t = y[7], theta = y[8], ptheta1t = y[9], psi = y[10], ppsi1t = y[11];
            const double sin_theta = std::sin(theta), cos_theta = std::cos(theta);
            const double sin_psi = std::sin(psi), cos_psi = std::cos(psi);
            const double KIN2 = ((pphi1t - ppsi1t * cos_theta) * sin_psi + ptheta1t * sin_theta * cos_psi) *
                                ((pphi1t - ppsi1t * cos_theta) * sin_psi + ptheta1t * sin_theta * cos_psi) /
                                (2.0 * IR[0] * sin_theta * sin_theta) +
                                ((pphi1t - ppsi1t * cos_theta) * cos_psi - ptheta1t * sin_theta * sin_psi) *
                                ((pphi1t - ppsi1t * cos_theta) * cos_psi - ptheta1t * sin_theta * sin_psi) /
                                (2.0 * IR[1] * sin_theta * sin_theta) +
                                ppsi1t * ppsi1t / (2.0 * IR[2]);
            const double sin_theta2t = std::sin(y[14]);
            const double KIN3 = y[15] * y[15] / (2.0 * I) + y[13] * y[13] / (2.0 * I * sin_theta2t * sin_theta2t);
            return k1 + KIN2 + KIN3;
        }
        return k1;
    }
};

// generate_normal, src/hawaii.c:1206-1230: cosine branch of Box-Muller, retry on non-finite log U
double normal(Mt& mt)
{
    for (;;) {
        const double U = mt.drand();
        const double V = mt.drand();
        const double logU = std::log(U);
        if (std::isnan(logU) || std::isinf(logU)) continue;
        return 1.0 * std::sqrt(-2 * logU) * std::cos(2.0 * M_PI * V);
    }
}

// q_generator, src/hawaii.c:1232-1290
void gen_q(const Sys& s, const hh_sampler& p, Mt& mt, double* y)
{
    const double RMIN = p.Rmin, RMAX = p.Rmax;
    const double R3 = mt.drand() * (RMAX * RMAX * RMAX - RMIN * RMIN * RMIN) + RMIN * RMIN * RMIN;
    y[4] = std::pow(R3, 1.0 / 3.0);
    y[0] = mt.drand() * 2.0 * M_PI;
    y[2] = std::acos(2.0 * mt.drand() - 1.0);
    if (s.linear()) {
        y[6] = mt.drand() * 2.0 * M_PI;
        y[8] = std::acos(2.0 * mt.drand() - 1.0);
    }
    if (s.rotor_linear()) {   // ROTOR: phi, theta, psi; then LINEAR_MOLECULE: phi, theta
        y[6] = mt.drand() * 2.0 * M_PI;
        y[8] = std::acos(2.0 * mt.drand() - 1.0);
        y[10] = mt.drand() * 2.0 * M_PI;
        y[12] = mt.drand() * 2.0 * M_PI;
        y[14] = std::acos(2.0 * mt.drand() - 1.0);
    }
}

// p_generator_linear_molecule, src/hawaii.c:1292-1303 (m: phi, pphi, theta, ptheta)
void gen_p_linear(double I, double T, Mt& mt, double* m)
{
    const double sqrt_IIKT = std::sqrt(I / kHkT * T);
    const double sin_theta = std::sin(m[2]);
    const double x0 = normal(mt), x1 = normal(mt);
    m[1] = sqrt_IIKT * sin_theta * x1;
    m[3] = sqrt_IIKT * x0;
}

// p_generator_rotor, src/hawaii.c:1305-1327 (m: phi, pphi, theta, ptheta, psi, ppsi)
void gen_p_rotor(const double* II, double T, Mt& mt, double* m)
{
    const double theta = m[2], psi = m[4];
    const double sqrt_IIx = std::sqrt(II[0] / kHkT * T), sqrt_IIy = std::sqrt(II[1] / kHkT * T);
    const double sqrt_IIz = std::sqrt(II[2] / kHkT * T);
    // the reference calls sincos, whose cosine can differ from cos in the last bit
    double sin_theta, cos_theta, sin_psi, cos_psi;
    ::sincos(theta, &sin_theta, &cos_theta);
    ::sincos(psi, &sin_psi, &cos_psi);
    const double x0 = normal(mt), x1 = normal(mt), x2 = normal(mt);
    m[1] = sqrt_IIx * sin_theta * sin_psi * x0 + sqrt_IIy * sin_theta * cos_psi * x1 + sqrt_IIz * cos_theta * x2;
    m[3] = sqrt_IIx * cos_psi * x0 - sqrt_IIy * sin_psi * x1;
    m[5] = sqrt_IIz * x2;
}

// p_generator + p_generator_linear_molecule, src/hawaii.c:1292-1380
void gen_p(const Sys& s, double T, Mt& mt, double* y)
{
    const double sqrt_MUKT = std::sqrt(s.mu / kHkT * T);
    const double sinTheta = std::sin(y[2]);
    const double x0 = normal(mt), x1 = normal(mt), x2 = normal(mt);
    y[1] = sqrt_MUKT * y[4] * sinTheta * x0;
    y[3] = sqrt_MUKT * y[4] * x1;
    y[5] = sqrt_MUKT * x2;
    if (s.linear()) gen_p_linear(s.I, T, mt, y + 6);
    if (s.rotor_linear()) {
        gen_p_rotor(s.IR, T, mt, y + 6);
        gen_p_linear(s.I, T, mt, y + 12);
    }
}

// reject, src/hawaii.c:1382-1421 (true = rejected)
bool rejected(const Sys& s, const double* y, double T, double pesmin, Mt& mt)
{
    const double u = mt.drand();
    const double Mval = std::exp(-pesmin * kHkT / T);
    const double M = 1.5 * Mval;
    const double kin_part = s.K(y);
    const double pot_part = s.V(y);
    const double proposal = std::exp(-kin_part * kHkT / T);
    const double desired = std::exp(-(kin_part + pot_part) * kHkT / T);
    const double weight = desired / proposal / M;
    return weight < u;
}

// the built-in surfaces in the plugin ABI, q = [Phi, Theta, R, monomer angles]
double hear_pes_lab(double* q) { return hear_V(q[2]); }
double co2ar_pes_lab(double* q) { return co2ar_V(q); }
double ch4co2_pes_lab(double* q) { return ch4co2_V(q); }

}  // namespace

extern "C" {

hh_pes_fn hh_host_pes(int system)
{
    switch (system) {
    case HH_SYS_HEAR: return hear_pes_lab;
    case HH_SYS_CO2AR: return co2ar_pes_lab;
    case HH_SYS_CH4CO2: return ch4co2_pes_lab;
    default: return nullptr;
    }
}

void hh_mt_fill(uint32_t seed, uint32_t* out, size_t n)
{
    Mt mt(seed);
    for (size_t i = 0; i < n; ++i) out[i] = mt.next();
}

__attribute__((visibility("hidden"))) long hh_sample_ics_impl(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long skip,
                                                              long ntraj, hh_pes_fn pes, double* ic, long* n_draws)
{
    const Sys s{sys->system, sys->mu, sys->I, sys->I_rotor, pes};
    const int qp = s.qp();
    Mt mt(seed);
    for (long i = 0; i < skip; ++i) (void)mt.drand();   // one word per draw: the stream after `skip` draws
    long accepted = 0, candidates = 0;
    double y[16];
    while (accepted < ntraj) {
        std::memset(y, 0, sizeof(y));
        gen_q(s, *smp, mt, y);
        if (smp->mode == HH_IC_R0_INJECT) {
            y[4] = smp->R0;
            gen_p(s, smp->T, mt, y);
            if (y[5] > 0.0) y[5] = -y[5];
            ++candidates;
        } else if (smp->mode == HH_IC_PHASE_SPACE) {
            gen_p(s, smp->T, mt, y);
            ++candidates;
        } else {
            gen_p(s, smp->T, mt, y);
            const double energy = s.K(y) + s.V(y);   // Hamiltonian = K + V (src/hawaii.c:1190-1204)
            ++candidates;
            if (rejected(s, y, smp->T, smp->pesmin, mt)) continue;
            if (smp->pair_state == HH_PAIR_FREE && energy < 0.0) continue;
            if (smp->pair_state == HH_PAIR_BOUND && energy > 0.0) continue;
        }
        std::memcpy(ic + accepted * qp, y, sizeof(double) * qp);
        ++accepted;
    }
    if (n_draws) *n_draws = mt.draws;
    return candidates;
}

}  // extern "C"
