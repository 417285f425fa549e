// chorin::div, grad, vgrad, flux and rhs (Chorin.hpp:19-70) on libxg.so.  The
// signatures are the reference's; the superedges, coordinates and boundary
// faces were uploaded once per mesh by xg_chocg_upload (ChoCG::setup), so the
// bodies pass the fields, and g_cfg's chorin settings come from
// xg_shim::chorin_cfg.  The source term of chorin::rhs (Chorin.cpp:985-1008)
// is the host's problems::SRC() evaluated at the nodes, applied after the
// device's advection terms, so every problem keeps its own source.
#include <string>
#include <vector>

#include "Chorin.hpp"
#include "Problems.hpp"
#include "xg_shim.hpp"

namespace xg_shim {

xg_chocfg chorin_cfg{0, 1, 0, 0.0, 0.0, 0.0};

}  // namespace xg_shim

namespace chorin {

namespace {

int comps(const tk::Fields& F) { return static_cast<int>(F.nprop()); }

}  // namespace

void div(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
         const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
         const std::array<std::vector<tk::real>, 3>& /*coord*/,
         const std::vector<std::size_t>& /*triinpoel*/, tk::real dt,
         const std::vector<tk::real>& P, const tk::Fields& G, const tk::Fields& U,
         std::vector<tk::real>& D, bool stab) {
  xg_ctx* c = xg_shim::t_ctx;
  if (U.nprop() < 3 || D.size() != U.nunk())
    Throw("xg: chorin::div: U needs 3 velocity components and D one value per node");
  if (stab && (P.size() != U.nunk() || G.nunk() != U.nunk() || G.nprop() < 3))
    Throw("xg: chorin::div: P and G need one value (G three) per node");
  xg_shim::check(c, xg_chorin_div(c, 0, dt, stab ? P.data() : nullptr,
                                  stab ? G.vec().data() : nullptr, stab ? comps(G) : 3,
                                  U.vec().data(), comps(U), D.data(), stab ? 1 : 0));
}

void grad(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
          const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
          const std::array<std::vector<tk::real>, 3>& /*coord*/,
          const std::vector<std::size_t>& /*triinpoel*/, const std::vector<tk::real>& U,
          tk::Fields& G) {
  xg_ctx* c = xg_shim::t_ctx;
  if (G.nunk() != U.size() || G.nprop() < 3)
    Throw("xg: chorin::grad: G needs three values per node of U");
  xg_shim::check(c, xg_chorin_grad(c, 0, U.data(), G.vec().data(), comps(G)));
}

void vgrad(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
           const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
           const std::array<std::vector<tk::real>, 3>& /*coord*/,
           const std::vector<std::size_t>& /*triinpoel*/, const tk::Fields& U,
           tk::Fields& G) {
  xg_ctx* c = xg_shim::t_ctx;
  if (G.nunk() != U.nunk() || G.nprop() != 3 * U.nprop())
    Throw("xg: chorin::vgrad: G must hold 3 gradients per component of U");
  xg_shim::check(c, xg_chorin_vgrad(c, 0, comps(U), U.vec().data(), G.vec().data()));
}

void flux(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
          const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
          const std::array<std::vector<tk::real>, 3>& /*coord*/,
          const std::vector<std::size_t>& /*triinpoel*/, const tk::Fields& U,
          const tk::Fields& G, tk::Fields& F) {
  xg_ctx* c = xg_shim::t_ctx;
  if (U.nprop() < 3 || G.nunk() != U.nunk() || G.nprop() < 9 || F.nunk() != U.nunk() ||
      F.nprop() != 3)
    Throw("xg: chorin::flux: U, G and F sizes disagree");
  xg_shim::check(c, xg_chorin_flux(c, 0, xg_shim::chorin_cfg.mu, U.vec().data(), comps(U),
                                   G.vec().data(), comps(G), F.vec().data()));
}

void rhs(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
         const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
         const std::array<std::vector<tk::real>, 3>& coord,
         const std::vector<std::size_t>& /*triinpoel*/, const std::vector<tk::real>& v,
         tk::real t, const std::vector<tk::real>& P, const tk::Fields& U,
         const tk::Fields& G, tk::Fields& R) {
  xg_ctx* c = xg_shim::t_ctx;
  const std::size_t n = U.nunk();
  if (R.nunk() != n || R.nprop() != U.nprop() || P.size() != n ||
      (xg_shim::chorin_cfg.flux == 1 && (G.nunk() != n || G.nprop() != 3 * U.nprop())))
    Throw("xg: chorin::rhs: P, U, G and R sizes disagree");
  xg_shim::check(c, xg_chorin_rhs(c, 0, &xg_shim::chorin_cfg, comps(U), P.data(),
                                  U.vec().data(), G.nunk() ? G.vec().data() : nullptr,
                                  R.vec().data()));
  auto src = problems::SRC();
  if (!src) return;
  if (v.size() != n || coord[0].size() != n || coord[1].size() != n || coord[2].size() != n)
    Throw("xg: chorin::rhs: the source term needs v and coord at every node");
  for (std::size_t p = 0; p < n; ++p) {
    auto s = src(coord[0][p], coord[1][p], coord[2][p], t, 0);
    if (s.size() > R.nprop()) Throw("xg: chorin::rhs: the source has more components than U");
    for (std::size_t k = 0; k < s.size(); ++k) R(p, k) -= s[k] * v[p];
  }
}

}  // namespace chorin
