/* host_fatal.h -- fatal errors of the library's host code.
 *
 * The reference stops on a bad argument with abort() (gmm.c:1514,
 * fisher.c:575).  Here the reason is printed first, through the VLFeat printf
 * hook and on stderr, both flushed, so that it is not lost in a buffer when
 * the process ends.
 */
#ifndef VL_HOST_FATAL_H
#define VL_HOST_FATAL_H

__attribute__((visibility("hidden"), noreturn, format(printf, 1, 2)))
void vl_host_fatal (char const * format, ...) ;

#endif
