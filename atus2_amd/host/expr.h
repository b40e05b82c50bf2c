// expr.h -- expression evaluator for the XML formulae (POTENTIAL_nD, GUESS_nD,
// AMP_T), covering the muParser syntax the reference's files use
// (ParameterHandler.cpp:298-312): numbers, the constants _pi and _e, XML
// CONSTANTS, bound variables (x, y, z, t), + - * / ^, unary + and -,
// comparisons, && ||, ?:, and the functions sin cos tan asin acos atan sinh
// cosh tanh asinh acosh atanh exp log ln log2 log10 sqrt abs rint sign
// min max sum avg, plus Heaviside and rect with the reference's copysign
// semantics (misc.cpp:40-53): sign(x) = copysign(1,x), Heaviside(x) =
// (1+sign(x))/2, rect(x,a,b) = Heaviside(x-a) - Heaviside(x-b).
//
// Precedence (lowest to highest): ?:, ||, &&, comparisons, + -, * /,
// unary + -, ^ (right associative).  So -x^2 = -(x^2) and 2^-1 = 0.5.
#ifndef ATUS_AMD_EXPR_H
#define ATUS_AMD_EXPR_H

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace atus_host {

// sign/Heaviside/rect of the reference (misc.cpp:40-53)
double sign_fn(double v);
double heaviside_fn(double v);
double rect_fn(double x, double xl, double xr);

class Expr {
 public:
  Expr();
  ~Expr();
  Expr(const Expr&);
  Expr& operator=(const Expr&);

  void DefineConst(const std::string& name, double value);
  void ClearConst();
  // binds a variable name to storage read at every Eval()
  void DefineVar(const std::string& name, const double* ptr);
  // sets the formula; it is parsed at the first Eval(), which throws
  // std::string with position information on errors
  void SetExpr(const std::string& text);
  double Eval() const;
  const std::string& GetExpr() const { return m_text; }

  struct Node;

 private:
  std::string m_text;
  std::map<std::string, double> m_consts;
  std::map<std::string, const double*> m_vars;
  std::shared_ptr<Node> m_root;
};

}  // namespace atus_host

#endif
