// no_device.cpp -- the library's device-side entry points for the host-only sanitizer
// builds: each one reports FGSS_ERR_DEVICE, so that the command-line tool runs its
// argument handling, input parsing and .bin writer up to its first device call.
// Collections in these builds are never uploaded (their device part stays null).
#include "fgss_internal.h"

static int no_device()
{
    return fgss_fail(FGSS_ERR_DEVICE, "host-only sanitizer build: no device code");
}

void fgss_device_release(fgss_collection *c)
{
    (void)c;   // nothing was uploaded
}

int fgss_collection_upload(fgss_collection *c)
{
    (void)c;
    return no_device();
}

int fgss_collection_from_text_device(const char *text, int64_t nbytes, int q, fgss_collection **out)
{
    (void)text;
    (void)nbytes;
    (void)q;
    if (out) *out = nullptr;
    return no_device();
}

int fgss_selfjoin(fgss_collection *c, float threshold, const fgss_join_opts *opts, fgss_pairs **out)
{
    (void)c;
    (void)threshold;
    (void)opts;
    if (out) *out = nullptr;
    return no_device();
}

void fgss_pairs_free(fgss_pairs *p)
{
    (void)p;   // no join ever returns pairs here
}

int fgss_pairs_print(const fgss_pairs *p, void *file)
{
    (void)p;
    (void)file;
    return no_device();
}

int fgss_collection_print_index(fgss_collection *c, float threshold, void *file)
{
    (void)c;
    (void)threshold;
    (void)file;
    return no_device();
}
