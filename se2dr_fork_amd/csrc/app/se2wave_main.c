/* se2wave_main.c -- the se2wave executable (se2wave.c:4916-4951). */
#include "se2wave.h"

int main(int argc, char **argv) { return se2w_main(argc, argv) ? 1 : 0; }
