// Units.hpp -- physical <-> lattice unit conversion.  Same derivation chain
// as the reference's UnitConverter (src/LBM/Units.cpp:20-199), because omega
// and the lattice force it yields must be the same double values:
//   dx = max_L / max_N;  nu_phys from units.nu_phys or U*L/Re;
//   u_lat from lbm.tau ((tau-0.5)*cs2*U*dx/nu_phys), lbm.u_lat, or 0.1;
//   dt = u_lat*dx/U;  nu_lat = nu_phys*dt/(dx*dx);  tau = 0.5 + nu_lat/cs2.
#pragma once

#include <optional>
#include <string>
#include <vector>

namespace muphfasa {

namespace unit_limits {
constexpr double tau_hard_lo = 0.5, tau_hard_hi = 2.0;
constexpr double tau_ideal_lo = 0.6, tau_ideal_hi = 1.5;
constexpr double u_lat_hard_hi = 0.3, u_lat_ideal_hi = 0.15;
constexpr double u_lat_default = 0.1;
}  // namespace unit_limits

class UnitConverter {
public:
    // inspect_only: skip the hard stability checks (diagnostic use).
    UnitConverter(double L_ref, double U_ref, double rho_ref, std::vector<int> N,
                  std::vector<double> domLo, std::vector<double> domHi, double cs2,
                  std::optional<double> Re = std::nullopt,
                  std::optional<double> nu_phys = std::nullopt,
                  std::optional<double> u_lat = std::nullopt,
                  std::optional<double> tau = std::nullopt, bool inspect_only = false,
                  bool verbose = false);

    double toLatticeLength(double x) const { return x / m_dx; }
    double toLatticeVelocity(double u) const { return u * m_dt / m_dx; }
    double toLatticeTime(double t) const { return t / m_dt; }
    double toLatticeViscosity(double nu) const { return nu * m_dt / (m_dx * m_dx); }
    double toLatticeAcceleration(double a) const { return a * m_dt * m_dt / m_dx; }
    double toLatticeDensity(double rho) const { return rho / m_rho_ref; }

    double toPhysicalLength(double x) const { return x * m_dx; }
    double toPhysicalVelocity(double u) const { return u * m_dx / m_dt; }
    double toPhysicalTime(double t) const { return t * m_dt; }
    double toPhysicalViscosity(double nu) const { return nu * m_dx * m_dx / m_dt; }
    double toPhysicalAcceleration(double a) const { return a * m_dx / (m_dt * m_dt); }
    double toPhysicalDensity(double rho) const { return rho * m_rho_ref; }

    double machNumber() const;
    double reynoldsNumber() const { return m_Re; }
    double relaxationTime() const { return m_tau; }
    double refLength() const { return m_L_ref; }
    double refVelocity() const { return m_U_ref; }
    double refDensity() const { return m_rho_ref; }
    double physViscosity() const { return m_nu_phys; }
    double latticeVelocity() const { return m_u_lat; }
    double latticeViscosity() const { return m_nu_lat; }
    double domainLength() const { return m_L_dom; }
    double gridCells() const { return m_N_grid; }
    double dx() const { return m_dx; }
    double dt() const { return m_dt; }
    std::string summary() const;

private:
    double m_L_ref, m_U_ref, m_Re = 0, m_nu_phys = 0, m_rho_ref;
    double m_L_dom = 0, m_N_grid = 0, m_cs2;
    double m_dx = 0, m_dt = 0, m_u_lat = 0, m_nu_lat = 0, m_tau = 0;
};

}  // namespace muphfasa
