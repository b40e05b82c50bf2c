// muParser.h -- the muParser 2.2 names the reference's drivers use
// (rt_solver.cpp:79-106, sob_solver.cpp:73-100, 117-138): mu::Parser with
// SetExpr / DefineVar / DefineConst / Eval, and mu::Parser::exception_type
// with GetMsg / GetExpr / GetToken / GetPos / GetCode, mapped onto this
// project's evaluator atus_host::Expr (the formulae of the XML files; its
// grammar is listed in expr.h).  Parse and evaluation errors are thrown as
// mu::Parser::exception_type, as muParser throws them.
#ifndef ATUS_AMD_MUPARSER_H
#define ATUS_AMD_MUPARSER_H

#include <string>

#include "expr.h"

namespace mu {

typedef double value_type;
typedef std::string string_type;

class ParserError {
 public:
  ParserError() = default;
  ParserError(const std::string& msg, const std::string& expr) : m_msg(msg), m_expr(expr) {}
  const std::string& GetMsg() const { return m_msg; }
  const std::string& GetExpr() const { return m_expr; }
  // the evaluator reports the position inside its message
  const std::string& GetToken() const { return m_token; }
  int GetPos() const { return -1; }
  int GetCode() const { return m_msg.empty() ? 0 : 1; }

 private:
  std::string m_msg, m_expr, m_token;
};

class Parser : public atus_host::Expr {
 public:
  typedef ParserError exception_type;

  void SetExpr(const std::string& text) { atus_host::Expr::SetExpr(text); }
  void DefineVar(const std::string& name, double* ptr) { atus_host::Expr::DefineVar(name, ptr); }
  void DefineConst(const std::string& name, double value) { atus_host::Expr::DefineConst(name, value); }
  double Eval() const {
    try {
      return atus_host::Expr::Eval();
    } catch (const std::string& s) {
      throw ParserError(s, GetExpr());
    }
  }
};

}  // namespace mu

#endif
