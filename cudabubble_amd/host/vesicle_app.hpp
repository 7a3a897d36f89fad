// vesicle_app.hpp -- C++ host side with the call shape of the reference's
// belastic (belastic.h:10-39): construct from an ini<i>.dat index, run the
// ETD driver (run_static_fEIF) or the Euler driver (run_static), with the
// reference's result folder, log lines and output files.  All computation
// goes through the C ABI of libvesicle.so (include/vesicle.h).
#pragma once
#include <string>

#include "vesicle.h"

namespace vbhost {

class belastic {
 public:
  // Reads <result_path>ini<folderindex>.dat (fluidbase.cpp:25-35), creates
  // <result_path><ifold>/ and copies the ini there (fluidbase.cpp:43-47),
  // continues a saved experiment if <folder><nx>x<ny>x<nz>.info exists
  // (fluidbase.cpp:2797-2825), else builds the initial shape of ninit.
  explicit belastic(int folderindex, const std::string& result_path = "./results/");
  ~belastic();
  belastic(const belastic&) = delete;
  belastic& operator=(const belastic&) = delete;

  void run_static_fEIF();   // belastic.cpp:554-786
  void run_static();        // belastic.cpp:286-469

  // E, V, A, E_b of the current field (belastic.cpp:68-86)
  double energy(double* V = nullptr, double* A = nullptr, double* Eb = nullptr);
  int status() const { return status_; }   // vb_status of the last driver exit
  long long steps() const { return steps_; }

  long long max_steps = 0;        // stop after this many driver steps (0: no cap)
  double output_every_s = 20.0;   // outputall() cadence in wall seconds (belastic.cpp:759)

 private:
  void output(const char* s);                                   // fluidbase.cpp:204-222
  void output(const char* name, double x);                      // fluidbase.cpp:286-297
  void output(const char* name, double x1, double x2);          // fluidbase.cpp:299-310
  void output9(const char* name, const double* x);              // fluidbase.cpp:312-323
  void output_field(const double* phi);                         // fluidbase.cpp:224-265
  void outputall();                                             // belastic.cpp:147-171
  bool check_folder();                                          // fluidbase.cpp:2885-2903
  void header(bool fixed);
  bool load_initial();
  void run(int scheme_driver);

  std::string result_path_, folder_, inifile_;
  int folderindex_ = 0, ifold_ = 0, ninit_ = 0;
  vb_params p_{};
  vb_solver* s_ = nullptr;
  int status_ = VB_OK;
  long long steps_ = 0;
};

}  // namespace vbhost
