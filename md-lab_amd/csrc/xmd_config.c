/* Compile/link flags for drivers of the engine, the counterpart of the
 * reference's xmd_config (xmd_config.c:4-26):
 *   cc driver.c $(xmd_config --xmd-flags) $(xmd_config --xmd-libs) */
#include <stdio.h>
#include <string.h>

#ifndef XMD_PREFIX
#define XMD_PREFIX "."
#endif

int main(int argc, char *argv[]) {
  if (argc < 2) {
    printf("Incorrect usage\n");
    return 1;
  }
  if (!strcmp(argv[1], "--xmd-flags"))
    printf("-I%s/include ", XMD_PREFIX);
  else if (!strcmp(argv[1], "--xmd-libs"))
    printf("-L%s/md-lab_amd/lib -Wl,-rpath,%s/md-lab_amd/lib -lxmd_x -L/opt/rocm/lib -lamdhip64 -lm -lpthread",
           XMD_PREFIX, XMD_PREFIX);
  fflush(stdout);
  return 0;
}
