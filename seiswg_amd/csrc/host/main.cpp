// `seiswg <config.ini>` (main.rs:33-100).
#include <cstdio>

#include "../../../include/seiswg_host.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "Usage: %s <config.ini>\n", argv[0]);
    return 1;
  }
  return seiswg_run(argv[1]);
}
