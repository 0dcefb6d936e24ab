// dsl_cpp_generator.h -- the mi355x backend as a StarPlat compiler plugin.
//
// Drop-in for StarPlat's src/backends/backend_mi355x/: the same interface as
// the OMP backend's generator class (src/backends/backend_omp/
// dsl_cpp_generator.h:57-58, setFileName + generate), driven by the parser's
// main after SymbolTableBuilder::buildST (src/parser/lrparser.y:710-821). It
// walks frontEndContext.getFuncList() (the annotated AST) and, instead of
// printing one OpenMP loop per forall, recognises each function's whole shape
// and emits a call into the hand-written HIP kernel library behind
// include/starplat_amd.h:
//
//   do-while pull sum over nodes_to + property copy   -> spm_compute_pr
//   fixedPoint + <Min, True> reduction over neighbors -> spm_compute_sssp
//   triple forall with is_an_edge and a scalar +=     -> spm_compute_tc
//   iterateInBFS sigma sum + iterateInReverse         -> spm_compute_bc
//
// A function whose arithmetic differs from the recognised idiom is refused
// (generate() returns false with a message) rather than lowered to a kernel
// that computes something else. Output: ../graphcode/generated_mi355x/
// <dsl base name>.{h,cc}, with the OMP backend's entry signatures (type
// mapping backend_omp/dsl_cpp_generator.cpp:2875-3054, parameters
// :3056-3096), so a driver written for the OMP output links unchanged.
#ifndef SPMI355X_DSL_CPP_GENERATOR_H
#define SPMI355X_DSL_CPP_GENERATOR_H

#include <map>
#include <string>
#include <vector>

#include "../../ast/ASTNodeTypes.hpp"

namespace spmi355x {

// One function lowered to one kernel-library call.
struct Plan {
  std::string kind;                           // "pr", "sssp", "tc", "bc"
  Function* func = nullptr;
  std::map<std::string, std::string> roles;   // role -> parameter or variable name
  std::vector<std::string> notes;
  bool diff_updated = true;                   // PageRank: the loop body writes diff
};

class dsl_cpp_generator {
 public:
  dsl_cpp_generator();
  // Base name of the output files: the last path component of f, without a
  // ".sp" suffix (backend_omp/dsl_cpp_generator.cpp:3231-3244 takes the last
  // component).
  void setFileName(char* f);
  // Lowers every function of frontEndContext.getFuncList() and writes the
  // .h/.cc pair; false (with a message on stderr and in error()) if a
  // function has no lowering or the files cannot be written.
  bool generate();
  // Output directory (default ../graphcode/generated_mi355x, relative to
  // src/ where StarPlat runs).
  void setOutputDir(const std::string& dir) { outDir = dir; }
  const std::string& error() const { return err; }

  // The lowering of one function (exposed for tests); false if no idiom
  // matches, with the reason in error().
  bool lower(Function* f, Plan& plan);
  // The emitted text of the plans, as generate() writes it.
  std::string headerText(const std::vector<Plan>& plans) const;
  std::string bodyText(const std::vector<Plan>& plans) const;

 private:
  std::string fileName;  // output base name
  std::string srcName;   // the DSL file's own name (last path component)
  std::string outDir;
  std::string err;

  int recognizePR(Function* f, Plan& p);
  int recognizeSSSP(Function* f, Plan& p);
  int recognizeTC(Function* f, Plan& p);
  int recognizeBC(Function* f, Plan& p);
  int fail(const std::string& msg);
};

}  // namespace spmi355x

#endif
