// C++-linkage forms of the two whole-volume SRRF entry points.
//
// The reference declares both with C++ linkage (GradientCalc/GradientCalc.h:4-5)
// and its callers import them that way (multispot5_place_methods.cc:48-49,
// multispot5.cc:77), so their object files reference the mangled names
// _Z16SRRFGradientCalcPfPiS_ and _Z10SRRFGetMagv.  The library exports these
// beside the extern "C" names of include/rfba_srrf.h, so such a caller links
// against it with only its dllimport attribute removed.  Both forward to the
// C entry points, which are reached here through assembler labels (the C++
// and the C function share one source-level name).
extern "C" int rfba_c_SRRFGradientCalc(float *inImage, int *imageDim, float *outImage) __asm__("SRRFGradientCalc");
extern "C" int rfba_c_SRRFGetMag(void) __asm__("SRRFGetMag");

int SRRFGradientCalc(float *inImage, int *imageDim, float *outImage) {
    return rfba_c_SRRFGradientCalc(inImage, imageDim, outImage);
}

int SRRFGetMag() { return rfba_c_SRRFGetMag(); }
