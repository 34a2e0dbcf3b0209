/* Test infrastructure (oracle): NOT part of the shipped product.
 *
 * Definitions of the reference's COMMON blocks for ref_driver_shim, the driver
 * linked against the gfx950 drop-in (libunicorn_fshim.so) instead of the
 * flang-built CKRATES/SOLVESP/SOLVEH (which would otherwise define them).  The
 * executable exports them (-rdynamic) so that the shared library binds to them,
 * as a Fortran program linked with --export-dynamic does. */
#define REF_COMMON
#include "ref_commons.h"
