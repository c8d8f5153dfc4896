// cexpr.h -- complex-valued expressions of the angular frequency w, for
// user-defined materials (MATERIAL ... ENDMATERIAL sections of a
// .scuffgeo file, libs/libMatProp/UserDefinedMaterials.cc:145-320):
//   wp = 1.37e16;  gamma = 5.32e13;  Eps(w) = 1 - wp^2/(w*(w+i*gamma));
// Numbers, the constants defined before the expression, w (rad/s), i,
// pi, + - * / ^ (right-associative), unary signs, parentheses and the
// functions sqrt exp log sin cos tan sinh cosh tanh abs.
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <string>

namespace scuffx {

struct CExpr;
using CExprPtr = std::shared_ptr<const CExpr>;

// parses text with the named constants bound; "" on success, else the error
std::string parse_cexpr(const std::string &text, const std::map<std::string, double> &constants, CExprPtr *out);
// value at angular frequency w (rad/s)
std::complex<double> eval_cexpr(const CExpr &e, std::complex<double> w);

}  // namespace scuffx
