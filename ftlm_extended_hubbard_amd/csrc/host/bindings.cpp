// bindings.cpp -- pybind11 module _ftlm_core over the drop-in C++ API.
// HubbardParams / nsites / version_string keep the reference's binding (python/bindings.cpp:11-38):
// same class, constructor keywords, defaults and attributes.  The hot-path classes and functions
// below are additions over that minimal module.

#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <stdexcept>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/ftlm_thermo.hpp"
#include "ftlm/hip/device.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lanczos.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/symmetry/orbit_k_block_basis.hpp"
#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/version.hpp"

namespace py = pybind11;
using cvec = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;

namespace {

cvec apply_host(int dim, const cvec& x, const std::function<void(const std::complex<double>*, std::complex<double>*)>& f) {
  if (x.ndim() != 1 || x.shape(0) != dim) throw std::invalid_argument("x must have shape (dim,)");
  cvec y(dim);
  f(x.data(), y.mutable_data());
  return y;
}

}  // namespace

PYBIND11_MODULE(_ftlm_core, m) {
  m.doc() = "MI355X FTLM hot path: the reference's C++ API on the gfx950 library";

  py::class_<ftlm::HubbardParams>(m, "HubbardParams")
      .def(py::init<>())
      .def(py::init([](int Lx, int Ly, double t, double tp, double U, double V, double phi_x, double phi_y) {
             ftlm::HubbardParams p;
             p.Lx = Lx;
             p.Ly = Ly;
             p.t = t;
             p.tp = tp;
             p.U = U;
             p.V = V;
             p.phi_x = phi_x;
             p.phi_y = phi_y;
             return p;
           }),
           py::arg("Lx") = 2, py::arg("Ly") = 2, py::arg("t") = 1.0, py::arg("tp") = 0.0, py::arg("U") = 0.0,
           py::arg("V") = 0.0, py::arg("phi_x") = 0.0, py::arg("phi_y") = 0.0)
      .def_readwrite("Lx", &ftlm::HubbardParams::Lx)
      .def_readwrite("Ly", &ftlm::HubbardParams::Ly)
      .def_readwrite("t", &ftlm::HubbardParams::t)
      .def_readwrite("tp", &ftlm::HubbardParams::tp)
      .def_readwrite("U", &ftlm::HubbardParams::U)
      .def_readwrite("V", &ftlm::HubbardParams::V)
      .def_readwrite("phi_x", &ftlm::HubbardParams::phi_x)
      .def_readwrite("phi_y", &ftlm::HubbardParams::phi_y);

  m.def("version_string", &ftlm::version_string);
  m.def("nsites", &ftlm::nsites, py::arg("params"));

  // ---- hot path
  py::class_<ftlm::SpinfulHopping>(m, "SpinfulHopping")
      .def(py::init<>())
      .def_readwrite("from_", &ftlm::SpinfulHopping::from)
      .def_readwrite("to", &ftlm::SpinfulHopping::to)
      .def_readwrite("spin", &ftlm::SpinfulHopping::spin)
      .def_readwrite("coeff", &ftlm::SpinfulHopping::coeff);
  py::class_<ftlm::NearestPair>(m, "NearestPair")
      .def(py::init<>())
      .def_readwrite("i", &ftlm::NearestPair::i)
      .def_readwrite("j", &ftlm::NearestPair::j);
  py::class_<ftlm::RectLattice>(m, "RectLattice")
      .def(py::init([](int Lx, int Ly) { return ftlm::RectLattice{Lx, Ly}; }), py::arg("Lx") = 1, py::arg("Ly") = 1)
      .def_readwrite("Lx", &ftlm::RectLattice::Lx)
      .def_readwrite("Ly", &ftlm::RectLattice::Ly)
      .def("n_sites", &ftlm::RectLattice::n_sites);
  m.def("build_hubbard_geometry", [](const ftlm::HubbardParams& p, const ftlm::RectLattice& lat) {
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, lat, &h, &pr);
    return py::make_tuple(h, pr);
  });

  py::class_<ftlm::FockBasis>(m, "FockBasis")
      .def(py::init<int, int, int>(), py::arg("n_sites"), py::arg("n_up"), py::arg("n_down"))
      .def("n_sites", &ftlm::FockBasis::n_sites)
      .def("n_up", &ftlm::FockBasis::n_up)
      .def("n_down", &ftlm::FockBasis::n_down)
      .def("dim", &ftlm::FockBasis::dim)
      .def("up_mask", &ftlm::FockBasis::up_mask)
      .def("down_mask", &ftlm::FockBasis::down_mask)
      .def("index_of", &ftlm::FockBasis::index_of);

  m.def("apply_extended_hubbard",
        [](const ftlm::HubbardParams& p, const ftlm::FockBasis& b, const std::vector<ftlm::SpinfulHopping>& h,
           const std::vector<ftlm::NearestPair>& pr, const cvec& x) {
          return apply_host(b.dim(), x, [&](const std::complex<double>* xi, std::complex<double>* yo) {
            ftlm::apply_extended_hubbard(p, b, h, pr, xi, yo);
          });
        });

  py::class_<ftlm::LanczosExtrema>(m, "LanczosExtrema")
      .def_readonly("min_eval", &ftlm::LanczosExtrema::min_eval)
      .def_readonly("max_eval", &ftlm::LanczosExtrema::max_eval)
      .def_readonly("steps_used", &ftlm::LanczosExtrema::steps_used);
  py::class_<ftlm::FtlmParams>(m, "FtlmParams")
      .def(py::init<>())
      .def_readwrite("n_random", &ftlm::FtlmParams::n_random)
      .def_readwrite("lanczos_steps", &ftlm::FtlmParams::lanczos_steps)
      .def_readwrite("seed", &ftlm::FtlmParams::seed);

  py::class_<ftlm::hip::SectorOperator>(m, "SectorOperator")
      .def(py::init<const ftlm::HubbardParams&, const ftlm::FockBasis&, const std::vector<ftlm::SpinfulHopping>&,
                    const std::vector<ftlm::NearestPair>&>())
      .def("dim", &ftlm::hip::SectorOperator::dim)
      .def("apply", [](const ftlm::hip::SectorOperator& o, const cvec& x) { return apply_host(o.dim(), x, o.apply_fn()); })
      .def("lanczos_extrema", [](const ftlm::hip::SectorOperator& o, int max_steps, unsigned seed) {
        return ftlm::lanczos_extrema(o.dim(), o.apply_fn(), max_steps, seed);
      }, py::arg("max_steps"), py::arg("seed") = 1)
      .def("lanczos_tridiagonal", [](const ftlm::hip::SectorOperator& o, int max_steps, unsigned seed) {
        std::vector<double> a, b;
        const int used = ftlm::lanczos_tridiagonal(o.dim(), o.apply_fn(), max_steps, seed, &a, &b);
        return py::make_tuple(used, a, b);
      })
      .def("ftlm_log_partition_complex", [](const ftlm::hip::SectorOperator& o, double beta, const ftlm::FtlmParams& par) {
        return ftlm::ftlm_log_partition_complex(o.dim(), o.apply_fn(), beta, par);
      })
      .def("ftlm_log_partition_real", [](const ftlm::hip::SectorOperator& o, double beta, const ftlm::FtlmParams& par) {
        return ftlm::ftlm_log_partition_real(o.dim(), o.apply_real_fn(), beta, par);
      });

  m.def("tridiagonal_extrema", &ftlm::tridiagonal_extrema);

  // HubbardMomentumBlock counterpart: the Gram-subspace momentum block of `sector` (same lattice,
  // particle numbers and hoppings) at K, applied through the full sector on the device.
  py::class_<ftlm::hip::GramBlockOperator>(m, "GramBlockOperator")
      .def(py::init([](const ftlm::hip::SectorOperator& sector, const ftlm::HubbardParams& p, const ftlm::FockBasis& fb,
                       ftlm::symmetry::MomentumSector K, const std::vector<ftlm::SpinfulHopping>& h,
                       const std::vector<ftlm::NearestPair>& pr) {
             const ftlm::hip::KBlockOperator kb(p, fb, K, h, pr);
             return new ftlm::hip::GramBlockOperator(sector, kb);
           }),
           py::arg("sector"), py::arg("params"), py::arg("fb"), py::arg("K"), py::arg("hoppings"), py::arg("nn_pairs"))
      .def("dim", &ftlm::hip::GramBlockOperator::dim)
      .def("apply", [](const ftlm::hip::GramBlockOperator& o, const cvec& x) { return apply_host(o.dim(), x, o.apply_fn()); })
      .def("lanczos_extrema", [](const ftlm::hip::GramBlockOperator& o, int max_steps, unsigned seed) {
        return ftlm::lanczos_extrema(o.dim(), o.apply_fn(), max_steps, seed);
      }, py::arg("max_steps"), py::arg("seed") = 1)
      .def("ftlm_log_partition_complex", [](const ftlm::hip::GramBlockOperator& o, double beta,
                                             const ftlm::FtlmParams& par) {
        return ftlm::ftlm_log_partition_complex(o.dim(), o.apply_fn(), beta, par);
      });

  py::class_<ftlm::symmetry::MomentumSector>(m, "MomentumSector")
      .def(py::init([](int kx, int ky, int lx, int ly) { return ftlm::symmetry::MomentumSector{kx, ky, lx, ly}; }),
           py::arg("kx") = 0, py::arg("ky") = 0, py::arg("lx") = 4, py::arg("ly") = 4)
      .def_readwrite("kx", &ftlm::symmetry::MomentumSector::kx)
      .def_readwrite("ky", &ftlm::symmetry::MomentumSector::ky)
      .def_readwrite("lx", &ftlm::symmetry::MomentumSector::lx)
      .def_readwrite("ly", &ftlm::symmetry::MomentumSector::ly)
      .def("flat_index", &ftlm::symmetry::MomentumSector::flat_index);

  py::class_<ftlm::symmetry::OrbitKBlockBasis>(m, "OrbitKBlockBasis")
      .def(py::init<>())
      .def("build", &ftlm::symmetry::OrbitKBlockBasis::build, py::arg("fb"), py::arg("K"), py::arg("survive_tol") = 1e-12)
      .def("dim", &ftlm::symmetry::OrbitKBlockBasis::dim)
      .def("reps", [](const ftlm::symmetry::OrbitKBlockBasis& b) {
        std::vector<uint32_t> r;
        for (const auto& e : b.entries()) r.push_back(ftlm::symmetry::pack_raw_state(e.rep));
        return r;
      })
      .def("reduced_index_of_rep", [](const ftlm::symmetry::OrbitKBlockBasis& b, uint32_t rep) {
        return b.reduced_index_of_rep(ftlm::symmetry::unpack_raw_state(rep));
      })
      .def("map_raw_to_rep", [](const ftlm::symmetry::OrbitKBlockBasis& b, uint32_t s) {
        int r = -1, dx = 0, dy = 0;
        double eta = 0.0;
        const bool ok = b.map_raw_to_rep(ftlm::symmetry::unpack_raw_state(s), &r, &dx, &dy, &eta);
        return py::make_tuple(ok, r, dx, dy, eta);
      });

  m.def("apply_orbit_k_block",
        [](const ftlm::HubbardParams& p, const ftlm::FockBasis& fb, const ftlm::symmetry::OrbitKBlockBasis& b,
           const std::vector<ftlm::SpinfulHopping>& h, const std::vector<ftlm::NearestPair>& pr, const cvec& x) {
          return apply_host(b.dim(), x, [&](const std::complex<double>* xi, std::complex<double>* yo) {
            ftlm::symmetry::apply_orbit_k_block(p, fb, b, h, pr, xi, yo);
          });
        });
}
