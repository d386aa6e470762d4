// Units.cpp -- see Units.hpp; follows src/LBM/Units.cpp:20-183 of the reference.
#include "Units.hpp"

#include <cmath>
#include <cstdio>
#include <sstream>
#include <stdexcept>

namespace muphfasa {

static void require(bool ok, const std::string& msg)
{
    if (!ok) throw std::runtime_error(msg);
}

UnitConverter::UnitConverter(double L_ref, double U_ref, double rho_ref, std::vector<int> N,
                             std::vector<double> domLo, std::vector<double> domHi, double cs2,
                             std::optional<double> Re, std::optional<double> nu_phys,
                             std::optional<double> u_lat, std::optional<double> tau,
                             bool inspect_only, bool verbose)
    : m_L_ref(L_ref), m_U_ref(U_ref), m_rho_ref(rho_ref), m_cs2(cs2)
{
    // setGridResolution (src/LBM/Units.cpp:131-167)
    const int D = (int)N.size();
    require(D >= 2 && (int)domLo.size() >= D && (int)domHi.size() >= D,
            "UnitConverter: grid and domain need one entry per dimension");
    int max_N = N[0];
    for (int n : N) max_N = n > max_N ? n : max_N;
    double max_L = -1;
    std::vector<double> res(D);
    for (int i = 0; i < D; ++i) {
        const double len = domHi[i] - domLo[i];
        max_L = len > max_L ? len : max_L;
        res[i] = len / double(N[i]);
    }
    for (int i = 0; i < D; ++i)
        for (int j = i + 1; j < D; ++j)
            require(std::abs(res[i] - res[j]) < 1.0e-12, "Lattice must be square (L/N equal in all directions)");
    m_L_dom = max_L;
    m_N_grid = max_N;
    m_dx = max_L / double(max_N);

    if (nu_phys) {
        require(!Re, "UnitConverter: Both units.Re and units.nu_phys were specified. Specify only one "
                     "and let the other be derived.");
        m_nu_phys = *nu_phys;
        m_Re = m_U_ref * m_L_ref / m_nu_phys;
    } else if (Re) {
        m_Re = *Re;
        m_nu_phys = m_U_ref * m_L_ref / m_Re;
    } else {
        throw std::runtime_error("UnitConverter: Neither units.Re nor units.nu_phys were specified.");
    }

    if (tau) {
        require(!u_lat, "UnitConverter: lbm.u_lat and lbm.tau were both specified. Specify only one.");
        const double nu_lat_from_tau = (*tau - 0.5) * cs2;
        m_u_lat = nu_lat_from_tau * m_U_ref * m_dx / m_nu_phys;
    } else if (u_lat) {
        m_u_lat = *u_lat;
    } else {
        m_u_lat = unit_limits::u_lat_default;
    }

    m_dt = m_u_lat * m_dx / m_U_ref;
    m_nu_lat = m_nu_phys * m_dt / (m_dx * m_dx);
    m_tau = 0.5 + m_nu_lat / m_cs2;

    if (inspect_only) return;
    require(m_tau > unit_limits::tau_hard_lo && m_tau < unit_limits::tau_hard_hi,
            "tau is not within the stability range 0.5 < tau < 2.0. Alter lbm.tau or lbm.u_lat to fix.");
    require(m_u_lat < unit_limits::u_lat_hard_hi,
            "u_lat does not satisfy low Mach number limit. Alter lbm.u_lat to fix.");
    if (verbose) std::fputs(summary().c_str(), stdout);
}

double UnitConverter::machNumber() const { return m_u_lat / std::sqrt(m_cs2); }

std::string UnitConverter::summary() const
{
    std::ostringstream s;
    s.precision(10);
    if (m_tau > unit_limits::tau_ideal_hi || m_tau < unit_limits::tau_ideal_lo)
        s << "Warning: the value of tau: " << m_tau << " is outside ideal range of 0.6 < tau < 1.5\n";
    if (m_u_lat > unit_limits::u_lat_ideal_hi)
        s << "Warning: the value of u_lat: " << m_u_lat
          << " is close to violating low Mach number constraint\n";
    s << "Physical Inputs:\nRe: " << m_Re << "\nU_phys: " << m_U_ref << "\nL_phys: " << m_L_ref
      << "\nnu_phys: " << m_nu_phys << "\nL_domain: " << m_L_dom
      << "\nStarting Simulation with Lattice Units:\nNgrid: " << m_N_grid << "\nU_lat: " << m_u_lat
      << "\nL_lat: " << toLatticeLength(m_L_ref) << "\nnu_lat: " << m_nu_lat << "\ntau: " << m_tau
      << "\nMa: " << machNumber() << "\nC_x: " << m_dx << "\nC_t: " << m_dt << "\n";
    return s.str();
}

}  // namespace muphfasa
