// emc_run: the process boundary of the engine, `emc_run input.txt`, like the
// reference's `main input.txt` (src/main.cpp:3-61).  Decks are read relative to
// the working directory, the band and phonon tables from dirname(exe)/../input
// (src/mcmodel.cpp:3974-3980), and the data/ files are written under ./data
// (created when missing).
// Options (after the deck): --device N (default 0), --seed S (default 1).
#include <libgen.h>
#include <limits.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/emc.h"

static std::string exe_dir() {
  char buf[PATH_MAX];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (n <= 0) return ".";
  buf[n] = 0;
  return std::string(dirname(buf));
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.txt [--device N] [--seed S]\n", argv[0]);
    return 2;
  }
  emc_opts o;
  emc_default_opts(&o);
  o.device = 0;
  for (int a = 2; a + 1 < argc; a += 2) {
    if (!std::strcmp(argv[a], "--device")) o.device = std::atoi(argv[a + 1]);
    else if (!std::strcmp(argv[a], "--seed")) o.seed = std::strtoull(argv[a + 1], nullptr, 10);
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[a]);
      return 2;
    }
  }
  const std::string tables = exe_dir() + "/../input";
  o.table_dir = tables.c_str();
  std::printf("FileName: %s\n", argv[1]);
  (void)mkdir("data", 0755);  // the data/ files go here (an existing directory is kept)
  const int st = emc_run_file(argv[1], &o);
  if (st != EMC_OK) {
    std::fprintf(stderr, "emc_run: status %d\n", st);
    return 1;
  }
  return 0;
}
