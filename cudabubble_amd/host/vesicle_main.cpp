// vesicle_main.cpp -- command line of the reference's main.cpp:5-26:
//   cudabubble [folderindex] [--results DIR] [--euler] [--max-steps N]
//              [--output-every SECONDS]
// runs belastic(folderindex).run_static_fEIF() (or run_static() with --euler)
// on results/ini<folderindex>.dat.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "vesicle_app.hpp"

int main(int argc, char* argv[]) {
  int index = 0;
  std::string results = "./results/";
  bool euler = false;
  long long max_steps = 0;
  double every = 20.0;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--results") && i + 1 < argc) {
      results = argv[++i];
    } else if (!strcmp(argv[i], "--euler")) {
      euler = true;
    } else if (!strcmp(argv[i], "--max-steps") && i + 1 < argc) {
      max_steps = atoll(argv[++i]);
    } else if (!strcmp(argv[i], "--output-every") && i + 1 < argc) {
      every = atof(argv[++i]);
    } else if (!strcmp(argv[i], "--help")) {
      printf("usage: %s [folderindex] [--results DIR] [--euler] [--max-steps N] [--output-every S]\n", argv[0]);
      return 0;
    } else {
      index = atoi(argv[i]);
    }
  }
  printf("Elastic bending energy program!\n");
  try {
    vbhost::belastic t(index, results);
    t.max_steps = max_steps;
    t.output_every_s = every;
    if (euler) t.run_static();
    else t.run_static_fEIF();
    const int st = t.status();
    return (st == VB_OK || st == VB_FINISHED || st == VB_ERR_DT_UNDERFLOW) ? 0 : 2;
  } catch (const std::exception& e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
