/* oracle_llr_cli.c -- the oracle's llr driver (TEST INFRASTRUCTURE). */
#include "oracle_llr.h"

int main(int argc, char **argv) { return ogl_ref_main(argc, argv); }
