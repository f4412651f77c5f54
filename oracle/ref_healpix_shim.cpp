/*
 * ref_healpix_shim.cpp — C entry points over the reference's own HEALPix sources
 * (src/Healpix_2.15a/healpix_base.cc, compiled where they lie by oracle/Makefile).
 * TEST INFRASTRUCTURE ONLY: pins the oracle's NEST-scheme restatement.
 */
#include "src/Healpix_2.15a/healpix_base.h"

extern "C" void ref_pix2ang_z_phi(int order, long ipix, double* z, double* phi)
{
    Healpix_Base b(order, NEST);
    b.pix2ang_z_phi((int)ipix, *z, *phi);
}

extern "C" void ref_nest2xyf(int order, long ipix, int* ix, int* iy, int* face)
{
    Healpix_Base b(order, NEST);
    b.nest2xyf((int)ipix, *ix, *iy, *face);
}

extern "C" long ref_xyf2nest(int order, int ix, int iy, int face)
{
    Healpix_Base b(order, NEST);
    return b.xyf2nest(ix, iy, face);
}
