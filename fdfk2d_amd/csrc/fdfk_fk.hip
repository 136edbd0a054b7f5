// FK background on the GPU (FP64): placeholder until the propagator kernels land.
extern "C" int fdfk_compute_background(fdfk_solver *s, const fdfk_fk *fk)
{
    (void)s; (void)fk;
    return fail(FDFK_ENOSYS, "GPU FK background not built yet");
}
extern "C" void comput_background_wavefield_(const int32_t *, const int32_t *, const int32_t *,
    const int32_t *, const int32_t *, const int32_t *, const int32_t *, const int32_t *,
    const int32_t *, const float *, const float *, const float *, const float *, const double *,
    const float *, const float *, const double *, const double *, const float *, const float *,
    float *, float *, float *, float *, float *, float *, float *, float *, float *, float *,
    float *, float *)
{
    std::fprintf(stderr, "comput_background_wavefield (fdfk2d_amd): not built yet\n");
}
