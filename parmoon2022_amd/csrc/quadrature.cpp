// Quadrature rule data. The points and weights are the published rules the
// reference tabulates (Gauss-Legendre products, Dunavant-type triangle rules,
// Keast-type tetrahedron rules); the point order inside each rule is the
// reference's, because local matrices are summed over points in rule order
// (src/AssembleRoutines/LocalAssembling.C:919-969).
#include "quadrature.hpp"
#include <mutex>

namespace pmcd {

namespace {

QuadFormula make_line(QuadRule t, std::vector<double> w, std::vector<double> x) {
  QuadFormula q; q.type = t; q.dim = 1; q.w = std::move(w); q.xi = std::move(x);
  return q;
}

// tensor rule with the first coordinate running fastest
QuadFormula tensor2(QuadRule t, const double* x1, const double* w1, int n) {
  QuadFormula q; q.type = t; q.dim = 2;
  for(int j = 0; j < n; ++j)
    for(int k = 0; k < n; ++k) {
      q.w.push_back(w1[j] * w1[k]);
      q.xi.push_back(x1[k]);
      q.eta.push_back(x1[j]);
    }
  return q;
}

// QuadFormula.C:1996-2049: weight = w[i]*w[j]*w[k], xi <- k, eta <- j, zeta <- i
QuadFormula tensor3(QuadRule t, const double* x1, const double* w1, int n) {
  QuadFormula q; q.type = t; q.dim = 3;
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      for(int k = 0; k < n; ++k) {
        q.w.push_back(w1[i] * w1[j] * w1[k]);
        q.xi.push_back(x1[k]);
        q.eta.push_back(x1[j]);
        q.zeta.push_back(x1[i]);
      }
  return q;
}

std::vector<QuadFormula> build_all() {
  std::vector<QuadFormula> all;
  // --- lines on [-1,1] (QuadFormula.C:9-37)
  all.push_back(make_line(QuadRule::Gauss1Line, {2.0}, {0.0}));
  {
    std::vector<double> x = {0.21132486540518711774, 0.78867513459481288226};
    for(auto& v : x) v = 2 * v - 1;
    all.push_back(make_line(QuadRule::Gauss2Line, {1.0, 1.0}, x));
  }
  {
    std::vector<double> w = {0.2777777777777778, 0.4444444444444444, 0.2777777777777778};
    std::vector<double> x = {0.11270166537925831149, 0.5, 0.88729833462074168851};
    for(int i = 0; i < 3; ++i) { w[i] *= 2; x[i] = 2 * x[i] - 1; }
    all.push_back(make_line(QuadRule::Gauss3Line, w, x));
  }
  all.push_back(make_line(QuadRule::Gauss4Line,
      {0.347854845137453857373063949, 0.652145154862546142626936051,
       0.652145154862546142626936051, 0.347854845137453857373063949},
      {-0.86113631159405257522394649, -0.33998104358485626480266576,
        0.33998104358485626480266576,  0.86113631159405257522394649}));
  // --- triangles on {(0,0),(1,0),(0,1)}
  {
    QuadFormula q; q.type = QuadRule::Gauss3Tria; q.dim = 2;
    const double a = 0.797426985353087322398025276169754, b = 0.101286507323456338800987361915123;
    const double c = 0.059715871789769820459117580973106, d = 0.470142064105115089770441209513447;
    const double wa = 0.0629695902724135762978419727500906, wc = 0.0661970763942530903688246939165759;
    q.w = {0.1125, wa, wa, wa, wc, wc, wc};
    q.xi = {0.333333333333333333333333333333333, a, b, b, c, d, d};
    q.eta = {0.333333333333333333333333333333333, b, a, b, d, c, d};
    all.push_back(q);
  }
  {
    QuadFormula q; q.type = QuadRule::Degree9Tria; q.dim = 2;
    const double w0 = 0.0485678981413994169096209912536443, w1 = 0.0156673501135695352684274156436046,
                 w2 = 0.0389137705023871396583696781497019, w3 = 0.0398238694636051265164458871320226,
                 w4 = 0.0127888378293490156308393992794999, w5 = 0.0216417696886446886446886446886446;
    const double a1 = 0.489682519198737627783706924836192, b1 = 0.020634961602524744432586150327616;
    const double a2 = 0.437089591492936637269930364435354, b2 = 0.125820817014126725460139271129292;
    const double a3 = 0.188203535619032730240961280467335, b3 = 0.623592928761934539518077439065330;
    const double a4 = 0.0447295133944527098651065899662763, b4 = 0.910540973211094580269786820067447;
    const double p = 0.741198598784498020690079873523423, r = 0.0368384120547362836348175987833851,
                 s = 0.221962989160765695675102527693192;
    const double th = 0.333333333333333333333333333333333;
    q.w = {w0, w1, w1, w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w5, w5, w5};
    q.xi = {th, a1, a1, b1, a2, a2, b2, a3, a3, b3, a4, a4, b4, p, r, p, s, r, s};
    q.eta = {th, a1, b1, a1, a2, b2, a2, a3, b3, a3, a4, b4, a4, r, p, s, p, s, r};
    all.push_back(q);
  }
  // --- quadrilaterals on [-1,1]^2 (QuadFormula.C:597-689)
  {
    const double g = 0.57735026918962576450914878050195746;
    QuadFormula q; q.type = QuadRule::Gauss2Quad; q.dim = 2;
    q.w = {1, 1, 1, 1}; q.xi = {-g, g, -g, g}; q.eta = {-g, -g, g, g};
    all.push_back(q);
  }
  {
    const double g = 0.77459666924148337703585307995647992;
    const double wa = 0.30864197530864197530864197530864198, wb = 0.49382716049382716049382716049382717,
                 wc = 0.79012345679012345679012345679012346;
    QuadFormula q; q.type = QuadRule::Gauss3Quad; q.dim = 2;
    q.w = {wa, wb, wa, wb, wc, wb, wa, wb, wa};
    q.xi = {-g, 0., g, -g, 0., g, -g, 0., g};
    q.eta = {-g, -g, -g, 0., 0., 0., g, g, g};
    all.push_back(q);
  }
  {
    const double a = 0.86113631159405257522394648889280951, b = 0.33998104358485626480266575910324469;
    const double wa = 0.12100299328560200552121209737014755, wb = 0.22685185185185185185185185185185185,
                 wc = 0.42529330301069429077508419892614875;
    QuadFormula q; q.type = QuadRule::Gauss4Quad; q.dim = 2;
    q.w = {wa, wb, wb, wa, wb, wc, wc, wb, wb, wc, wc, wb, wa, wb, wb, wa};
    const double x1[4] = {-a, -b, b, a};
    for(int j = 0; j < 4; ++j)
      for(int k = 0; k < 4; ++k) { q.xi.push_back(x1[k]); q.eta.push_back(x1[j]); }
    all.push_back(q);
  }
  // --- tetrahedra on {(0,0,0),(1,0,0),(0,1,0),(0,0,1)} (QuadFormula.C:1678-1800)
  {
    const double a = 0.138196601125010515179541316563, b = 0.585410196624968454461376050311;
    QuadFormula q; q.type = QuadRule::P2Tetra; q.dim = 3;
    q.w = {1. / 24., 1. / 24., 1. / 24., 1. / 24.};
    q.xi = {a, a, a, b}; q.eta = {a, a, b, a}; q.zeta = {a, b, a, a};
    all.push_back(q);
  }
  {
    QuadFormula q; q.type = QuadRule::P4Tetra; q.dim = 3;
    const double w0 = -.0131555555555555555555555555556, w1 = .00762222222222222222222222222222,
                 w2 = .0248888888888888888888888888889;
    const double s = .0714285714285714285714285714286, t = .785714285714285714285714285714;
    const double u = .399403576166799204996102147462, v = .100596423833200795003897852538;
    q.w = {w0, w1, w1, w1, w1, w2, w2, w2, w2, w2, w2};
    q.xi = {.25, s, t, s, s, u, u, v, v, v, u};
    q.eta = {.25, s, s, t, s, u, v, u, v, u, v};
    q.zeta = {.25, s, s, s, t, v, u, u, u, v, v};
    all.push_back(q);
  }
  {
    QuadFormula q; q.type = QuadRule::P5Tetra; q.dim = 3;
    const double w0 = 0.602678571428571428571428571428571e-2, w1 = 0.302836780970891758063769725577305e-1,
                 w2 = 0.116452490860289694108936091443380e-1, w3 = 0.109491415613864593456430191124068e-1;
    const double th = 0.333333333333333333333333333333333;
    const double a = 0.909090909090909090909090909090909e-1, b = 0.727272727272727272727272727272727;
    const double c = 0.665501535736642982398804642263025e-1, d = 0.433449846426335701760119535773697;
    q.w = {w0, w0, w0, w0, w1, w2, w2, w2, w2, w3, w3, w3, w3, w3, w3};
    q.xi = {th, th, 0, th, 0.25, a, a, b, a, c, d, d, c, c, d};
    q.eta = {th, 0, th, th, 0.25, a, b, a, a, d, c, d, c, d, c};
    q.zeta = {0, th, th, th, 0.25, b, 0.90909090909090909090909090909091e-1,
              0.909090909090909090909090909090911e-1, 0.909090909090909090909090909090911e-1,
              0.433449846426335701760119535773699, 0.433449846426335701760119535773699,
              0.66550153573664298239880464226304e-1, 0.433449846426335701760119535773699,
              0.66550153573664298239880464226304e-1, 0.665501535736642982398804642263035e-1};
    all.push_back(q);
  }
  // --- hexahedra on [-1,1]^3 (QuadFormula.C:1981-2049)
  {
    const double g = 0.5773502691896257645091489;
    QuadFormula q; q.type = QuadRule::Gauss2Hexa; q.dim = 3;
    q.w.assign(8, 1.0);
    q.xi = {-g, g, -g, g, -g, g, -g, g};
    q.eta = {-g, -g, g, g, -g, -g, g, g};
    q.zeta = {-g, -g, -g, -g, g, g, g, g};
    all.push_back(q);
  }
  {
    const double w1[3] = {5. / 9., 8. / 9., 5. / 9.};
    const double x1[3] = {-0.7745966692414833770358530, 0, 0.7745966692414833770358530};
    all.push_back(tensor3(QuadRule::Gauss3Hexa, x1, w1, 3));
  }
  {
    const double w1[4] = {0.652145154862546142626936051, 0.347854845137453857373063949,
                          0.347854845137453857373063949, 0.652145154862546142626936051};
    const double x1[4] = {0.33998104358485626480266576, 0.86113631159405257522394649,
                          -0.86113631159405257522394649, -0.33998104358485626480266576};
    all.push_back(tensor3(QuadRule::Gauss4Hexa, x1, w1, 4));
  }
  (void)tensor2;
  return all;
}

}  // namespace

const QuadFormula& quad_formula(QuadRule r) {
  static std::vector<QuadFormula> all = build_all();
  for(auto& q : all) if(q.type == r) return q;
  PMCD_FAIL("unknown quadrature rule ", (int)r);
}

const char* quad_rule_name(QuadRule r) {
  switch(r) {
    case QuadRule::Gauss1Line: return "Gauss1Line";
    case QuadRule::Gauss2Line: return "Gauss2Line";
    case QuadRule::Gauss3Line: return "Gauss3Line";
    case QuadRule::Gauss4Line: return "Gauss4Line";
    case QuadRule::Gauss3Tria: return "Gauss3Tria";
    case QuadRule::Degree9Tria: return "Degree9Tria";
    case QuadRule::Gauss2Quad: return "Gauss2Quad";
    case QuadRule::Gauss3Quad: return "Gauss3Quad";
    case QuadRule::Gauss4Quad: return "Gauss4Quad";
    case QuadRule::P2Tetra: return "P2Tetra";
    case QuadRule::P4Tetra: return "P4Tetra";
    case QuadRule::P5Tetra: return "P5Tetra";
    case QuadRule::Gauss2Hexa: return "Gauss2Hexa";
    case QuadRule::Gauss3Hexa: return "Gauss3Hexa";
    case QuadRule::Gauss4Hexa: return "Gauss4Hexa";
  }
  return "?";
}

// QuadratureFormulaDatabase.C:29-131, restricted to degrees <= 6.
QuadRule cell_rule_for(Shape s, int p, bool linear_mode) {
  const int deg = (linear_mode ? 2 : 3) * p;
  PMCD_CHECK(deg >= 0 && deg <= 6, "unsupported quadrature degree ", deg);
  switch(s) {
    case Shape::Triangle: return deg <= 5 ? QuadRule::Gauss3Tria : QuadRule::Degree9Tria;
    case Shape::Quadrangle:
      return deg <= 2 ? QuadRule::Gauss2Quad : (deg <= 5 ? QuadRule::Gauss3Quad : QuadRule::Gauss4Quad);
    case Shape::Tetrahedron:
      return deg <= 2 ? QuadRule::P2Tetra : (deg <= 4 ? QuadRule::P4Tetra : QuadRule::P5Tetra);
    case Shape::Hexahedron:
      return deg <= 2 ? QuadRule::Gauss2Hexa : (deg <= 5 ? QuadRule::Gauss3Hexa : QuadRule::Gauss4Hexa);
  }
  PMCD_FAIL("bad shape");
}

QuadRule line_rule_for_degree(int deg) {
  if(deg <= 1) return QuadRule::Gauss1Line;
  if(deg <= 3) return QuadRule::Gauss2Line;
  if(deg <= 5) return QuadRule::Gauss3Line;
  if(deg <= 7) return QuadRule::Gauss4Line;
  PMCD_FAIL("unsupported line quadrature degree ", deg);
}

}  // namespace pmcd
