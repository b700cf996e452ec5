/* A C caller of the drop-in symbol, as the reference's cusrc/main.cu:6-10
 * calls it: green(iname, "out4.txt", "signal.txt") with paths relative to
 * the working directory.  Built against include/grtm.h only (the reference's
 * popm/popm.h pulls in the CUDA headers through mathematics.h:12-13; this
 * one-line include is the change a caller makes).  Exits with the status
 * green() recorded. */
#include <stdio.h>

#include "grtm.h"

int main(int argc, char **argv)
{
    char *iname = argc > 1 ? argv[1] : "input.txt";
    char *oname_w = argc > 2 ? argv[2] : "out4.txt";
    char *oname_t = argc > 3 ? argv[3] : "signal.txt";
    green(iname, oname_w, oname_t);
    const int st = grtm_last_status();
    if (st != GRTM_OK) fprintf(stderr, "green: %s\n", grtm_status_str(st));
    return st == GRTM_OK ? 0 : 1;
}
