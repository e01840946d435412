/* Test hook linked into the reference build under oracle/_ref (never into the
 * product).  When the environment variable BS_REF_PRINT is set, a constructor
 * switches on the reference's own dump flags before its main program runs:
 *   'b' -> print_basis   (module verbosity, src/bmodules_main.f90:126)
 *   'j' -> print_jumps   (src/bmodules_main.f90:144)
 *   's' -> print_sectors (src/bmodules_main.f90:119)
 * The dumps are then written by the reference's own writers
 * (src/boutputlib.f90:473-560, src/boutputlib.f90:1258-1351).
 * A Fortran default logical is a 4-byte integer; .true. is 1. */
#include <stdlib.h>
#include <string.h>

extern int _QMverbosityEprint_basis;
extern int _QMverbosityEprint_jumps;
extern int _QMverbosityEprint_sectors;

__attribute__((constructor)) static void bs_ref_hooks_init(void) {
  const char *s = getenv("BS_REF_PRINT");
  if (!s) return;
  if (strchr(s, 'b')) _QMverbosityEprint_basis = 1;
  if (strchr(s, 'j')) _QMverbosityEprint_jumps = 1;
  if (strchr(s, 's')) _QMverbosityEprint_sectors = 1;
}
