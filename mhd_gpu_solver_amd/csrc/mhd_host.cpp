// Host-side logic of the hip backend that needs no device: initial conditions and the solid
// wall-fill plan. Both run in fp64 on the CPU, as the reference does them on the host
// (include/initial_condition/inititial_condition.hpp:226-270, include/mhd/mask.hpp:85-149).
#include "mhd_host.hpp"

#include "mhd_hip.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace mhdhip {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Prim {
    double rho = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, p = 0.0, bx = 0.0, by = 0.0, bz = 0.0,
           psi = 0.0;
};

// to_conservative: include/mhd/mhd_physics.hpp:41-59
void to_cons(const Prim& W, double gamma, double (&S)[9])
{
    S[0] = W.rho;
    S[1] = W.rho * W.vx;
    S[2] = W.rho * W.vy;
    S[3] = W.rho * W.vz;
    S[5] = W.bx;
    S[6] = W.by;
    S[7] = W.bz;
    S[8] = W.psi;
    const double e_kin = 0.5 * W.rho * (W.vx * W.vx + W.vy * W.vy + W.vz * W.vz);
    const double e_mag = 0.5 * (W.bx * W.bx + W.by * W.by + W.bz * W.bz);
    S[4] = W.p / (gamma - 1.0) + e_kin + e_mag;
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// uniform in [-1, 1) from a 64-bit hash
double unit_sym(uint64_t h)
{
    return static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

enum class IC {
    Uniform, Sine, Blast, OrszagTang, Sod, BrioWu, Rotor, KH, Alfven,
    OrszagTangB0, RotorB0, KHTanh, AlfvenOblique
};

// include/initial_condition/inititial_condition.hpp:194-224 plus this backend's workloads
bool parse_ic(const std::string& s, IC& out)
{
    if (s == "uniform" || s == "Uniform") out = IC::Uniform;
    else if (s == "sine" || s == "SineWave") out = IC::Sine;
    else if (s == "blast" || s == "BlastWave") out = IC::Blast;
    else if (s == "orszag_tang" || s == "OrszagTang") out = IC::OrszagTang;
    else if (s == "sod" || s == "Sod") out = IC::Sod;
    else if (s == "brio_wu" || s == "BrioWu") out = IC::BrioWu;
    else if (s == "rotor" || s == "Rotor") out = IC::Rotor;
    else if (s == "kelvin_helmholtz" || s == "KelvinHelmholtz" || s == "kh") out = IC::KH;
    else if (s == "alfven" || s == "alfven_wave" || s == "AlfvenWave") out = IC::Alfven;
    else if (s == "orszag_tang_b0") out = IC::OrszagTangB0;
    else if (s == "rotor_b0") out = IC::RotorB0;
    else if (s == "kh_tanh") out = IC::KHTanh;
    else if (s == "alfven_oblique") out = IC::AlfvenOblique;
    else return false;
    return true;
}

// Rotor profile shared by "rotor" and "rotor_b0": inititial_condition.hpp:103-135
Prim rotor(double x, double y, double bscale)
{
    const double x0 = 0.5, y0 = 0.5, r0 = 0.1, r1 = 0.115;
    const double dx = x - x0;
    const double dy = y - y0;
    const double r = std::hypot(dx, dy);
    Prim W;
    W.bx = bscale * (5.0 / std::sqrt(4.0 * kPi));
    W.p = 1.0;
    double f = 0.0;
    if (r <= r0) {
        f = 1.0;
        W.rho = 10.0;
    } else if (r >= r1) {
        f = 0.0;
        W.rho = 1.0;
    } else {
        f = (r1 - r) / (r1 - r0);
        W.rho = 1.0 + 9.0 * f;
    }
    const double v0 = 2.0;
    if (r > 0.0) {
        W.vx = -f * v0 * dy / r0;
        W.vy = f * v0 * dx / r0;
    }
    return W;
}

Prim point(IC ic, double x, double y, double b0, uint64_t seed, uint64_t cell)
{
    Prim W;
    switch (ic) {
    case IC::Uniform:  // :30-33
        W.rho = 1.0;
        W.p = 1.0;
        break;
    case IC::Sine:  // :35-41
        W.rho = 1.0 + 0.1 * std::sin(2.0 * kPi * x);
        W.p = 1.0;
        break;
    case IC::Blast: {  // :44-56
        const double r = std::hypot(x - 0.5, y - 0.5);
        W.rho = 1.0;
        W.p = (r < 0.1) ? 10.0 : 0.1;
        W.bx = 1.0 / std::sqrt(2.0);
        W.by = 1.0 / std::sqrt(2.0);
        break;
    }
    case IC::OrszagTang:  // :58-69
        W.rho = 25.0 / (36.0 * kPi);
        W.vx = -std::sin(2.0 * kPi * y);
        W.vy = std::sin(2.0 * kPi * x);
        W.p = 5.0 / (12.0 * kPi);
        W.bx = -std::sin(2.0 * kPi * y);
        W.by = std::sin(4.0 * kPi * x);
        break;
    case IC::OrszagTangB0:
        W.rho = 25.0 / (36.0 * kPi);
        W.vx = -std::sin(2.0 * kPi * y);
        W.vy = std::sin(2.0 * kPi * x);
        W.p = 5.0 / (12.0 * kPi);
        W.bx = b0 * -std::sin(2.0 * kPi * y);
        W.by = b0 * std::sin(4.0 * kPi * x);
        break;
    case IC::Sod:  // :72-83
        if (x < 0.5) {
            W.rho = 1.0;
            W.p = 1.0;
        } else {
            W.rho = 0.125;
            W.p = 0.1;
        }
        break;
    case IC::BrioWu:  // :86-100
        W.bx = 0.75;
        if (x < 0.5) {
            W.rho = 1.0;
            W.p = 1.0;
            W.by = 1.0;
        } else {
            W.rho = 0.125;
            W.p = 0.1;
            W.by = -1.0;
        }
        break;
    case IC::Rotor:
        W = rotor(x, y, 1.0);
        break;
    case IC::RotorB0:
        W = rotor(x, y, b0);
        break;
    case IC::KH: {  // :138-149
        const bool in = (y > 0.25 && y < 0.75);
        W.rho = in ? 2.0 : 1.0;
        W.vx = in ? 0.5 : -0.5;
        W.vy = 0.01 * std::sin(2.0 * kPi * x);
        W.p = 2.5;
        break;
    }
    case IC::KHTanh: {
        // periodic double shear layer with tanh profiles of width 0.05
        const double w = 0.05;
        const double s = std::tanh((y - 0.25) / w) - std::tanh((y - 0.75) / w);  // ~2 inside
        W.rho = 1.5 + 0.5 * (s - 1.0);
        W.vx = 0.5 * (s - 1.0);
        W.vy = 1.0e-2 * unit_sym(splitmix64(seed ^ splitmix64(cell)));
        W.p = 2.5;
        W.bx = 0.1;
        break;
    }
    case IC::Alfven: {  // :153-172
        const double delta = 1.0e-3;
        const double rho = 1.0;
        const double phase = 2.0 * kPi * x;
        W.rho = rho;
        W.p = 0.1;
        W.bx = 1.0;
        W.by = delta * std::cos(phase);
        W.bz = delta * std::sin(phase);
        W.vy = -(delta / std::sqrt(rho)) * std::cos(phase);
        W.vz = -(delta / std::sqrt(rho)) * std::sin(phase);
        break;
    }
    case IC::AlfvenOblique: {
        // wave vector 2 pi (1, 2): angle atan(2), wavelength 1/sqrt(5), periodic on [0,1]^2
        const double amp = 0.1;
        const double ca = 1.0 / std::sqrt(5.0);
        const double sa = 2.0 / std::sqrt(5.0);
        const double phase = 2.0 * kPi * (x + 2.0 * y);
        const double b_par = 1.0;
        const double b_perp = amp * std::cos(phase);
        const double v_perp = -amp * std::cos(phase);
        W.rho = 1.0;
        W.p = 0.1;
        W.bx = b_par * ca - b_perp * sa;
        W.by = b_par * sa + b_perp * ca;
        W.bz = amp * std::sin(phase);
        W.vx = -v_perp * sa;
        W.vy = v_perp * ca;
        W.vz = -amp * std::sin(phase);
        break;
    }
    }
    return W;
}

}  // namespace

int init_ic(const char* name, const mhd_hip_grid& g, double gamma, double b0, uint64_t seed,
            double* const comps[9])
{
    IC ic;
    if (name == nullptr || !parse_ic(name, ic)) {
        throw std::invalid_argument(std::string("Unknown initial condition: ") +
                                    (name ? name : "(null)"));
    }
    const int nxt = g.nx + 2 * g.ng;
    const double dx = (g.x_max - g.x_min) / static_cast<double>(g.nx);
    const double dy = (g.y_max - g.y_min) / static_cast<double>(g.ny);
    for (int j = g.ng; j < g.ng + g.ny; ++j) {
        for (int i = g.ng; i < g.ng + g.nx; ++i) {
            // cell centres: include/grid/grid.hpp:38-45
            const double x = g.x_min + (i - g.ng + 0.5) * dx;
            const double y = g.y_min + (j - g.ng + 0.5) * dy;
            const uint64_t cell = static_cast<uint64_t>(j - g.ng) * g.nx + (i - g.ng);
            double S[9];
            to_cons(point(ic, x, y, b0, seed, cell), gamma, S);
            const size_t k = static_cast<size_t>(j) * nxt + i;
            for (int c = 0; c < 9; ++c) {
                comps[c][k] = S[c];
            }
        }
    }
    return 0;
}

// include/mhd/mask.hpp:97-146: first interior fluid neighbour in the order +x, -x, +y, -y
size_t wall_plan(const mhd_hip_grid& g, const uint8_t* mask, std::vector<long long>& plan)
{
    plan.clear();
    const int nxt = g.nx + 2 * g.ng;
    auto fluid = [&](int i, int j) {
        return mask[static_cast<size_t>(j - g.ng) * g.nx + (i - g.ng)] != 0;
    };
    for (int j = g.ng; j < g.ng + g.ny; ++j) {
        for (int i = g.ng; i < g.ng + g.nx; ++i) {
            if (fluid(i, j)) {
                continue;
            }
            const int cand[4][3] = {{i + 1, j, 0}, {i - 1, j, 0}, {i, j + 1, 1}, {i, j - 1, 1}};
            long long src = -1;
            int axis = 0;
            for (const auto& c : cand) {
                if (c[0] < g.ng || c[0] >= g.ng + g.nx || c[1] < g.ng || c[1] >= g.ng + g.ny) {
                    continue;
                }
                if (!fluid(c[0], c[1])) {
                    continue;
                }
                src = static_cast<long long>(c[1]) * nxt + c[0];
                axis = c[2];
                break;
            }
            plan.push_back(static_cast<long long>(j) * nxt + i);
            plan.push_back(src);
            plan.push_back(axis);
        }
    }
    return plan.size() / 3;
}

}  // namespace mhdhip
