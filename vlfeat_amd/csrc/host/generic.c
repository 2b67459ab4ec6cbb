/* generic.c -- runtime services: allocator and printf hooks, thread count,
 * SIMD switch, CPU feature queries, per-thread random generator, timers.
 *
 * Semantics follow vl/generic.c (VLFeat 0.9.21): the allocator hooks default
 * to the C library (generic.c:1289-1362), the printf hook to printf
 * (generic.c:1385-1400), SIMD is enabled by default (generic.c:1602-1632), and
 * the error code and message are per calling thread (generic.c:1238-1275),
 * every calling thread owns its own MT19937 generator seeded on first use by
 * vl_rand_init (generic.c:1460-1464, state created in generic.c:1512-1540:
 * vl_rand_init leaves the generator unseeded, so its first draw seeds it with
 * 5489).
 */
#include "vl/generic.h"
#include "vl/random.h"
#include "host_fatal.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

static void *(*g_malloc)(size_t) = malloc ;
static void *(*g_realloc)(void*, size_t) = realloc ;
static void *(*g_calloc)(size_t, size_t) = calloc ;
static void (*g_free)(void*) = free ;
static printf_func_t g_printf = printf ;
static vl_bool g_simd = VL_TRUE ;
static vl_size g_num_threads = 0 ;   /* 0 = all cores */

VL_EXPORT char const * vl_get_version_string (void) { return VL_VERSION_STRING " (MI355X)" ; }

VL_EXPORT void vl_set_simd_enabled (vl_bool x) { g_simd = x ; }
VL_EXPORT vl_bool vl_get_simd_enabled (void) { return g_simd ; }

VL_EXPORT vl_bool vl_cpu_has_avx (void) { return __builtin_cpu_supports("avx") ? VL_TRUE : VL_FALSE ; }
VL_EXPORT vl_bool vl_cpu_has_sse3 (void) { return __builtin_cpu_supports("sse3") ? VL_TRUE : VL_FALSE ; }
VL_EXPORT vl_bool vl_cpu_has_sse2 (void) { return __builtin_cpu_supports("sse2") ? VL_TRUE : VL_FALSE ; }

VL_EXPORT vl_size
vl_get_num_cpus (void)
{
  long n = sysconf(_SC_NPROCESSORS_ONLN) ;
  return n > 0 ? (vl_size)n : 1 ;
}

VL_EXPORT vl_size vl_get_thread_limit (void) { return vl_get_num_cpus() ; }

VL_EXPORT vl_size
vl_get_max_threads (void)
{
  return g_num_threads ? g_num_threads : vl_get_num_cpus() ;
}

VL_EXPORT void vl_set_num_threads (vl_size n) { g_num_threads = n ; }

VL_EXPORT void
vl_set_alloc_func (void *(*malloc_func) (size_t),
                   void *(*realloc_func) (void*,size_t),
                   void *(*calloc_func) (size_t, size_t),
                   void  (*free_func) (void*))
{
  g_malloc = malloc_func ;
  g_realloc = realloc_func ;
  g_calloc = calloc_func ;
  g_free = free_func ;
}

VL_EXPORT void * vl_malloc (size_t n) { return g_malloc(n) ; }
VL_EXPORT void * vl_realloc (void * ptr, size_t n) { return g_realloc(ptr, n) ; }
VL_EXPORT void * vl_calloc (size_t n, size_t size) { return g_calloc(n, size) ; }
VL_EXPORT void * vl_memalign (size_t n, size_t size) { (void)size ; return g_malloc(n) ; }
VL_EXPORT void vl_free (void * ptr) { g_free(ptr) ; }

static int
quiet_printf (char const * format, ...)
{
  (void)format ;
  return 0 ;
}

VL_EXPORT void vl_set_printf_func (printf_func_t f) { g_printf = f ? f : quiet_printf ; }
VL_EXPORT printf_func_t vl_get_printf_func (void) { return g_printf ; }

/* ---- per-thread random generator ---------------------------------------- */

static pthread_key_t g_rand_key ;
static pthread_once_t g_rand_once = PTHREAD_ONCE_INIT ;

static void rand_destroy (void * p) { free(p) ; }
static void rand_key_create (void) { pthread_key_create(&g_rand_key, rand_destroy) ; }

VL_EXPORT VlRand *
vl_get_rand (void)
{
  VlRand * r ;
  pthread_once(&g_rand_once, rand_key_create) ;
  r = (VlRand*) pthread_getspecific(g_rand_key) ;
  if (!r) {
    r = (VlRand*) malloc(sizeof(VlRand)) ;
    vl_rand_init(r) ;
    pthread_setspecific(g_rand_key, r) ;
  }
  return r ;
}

/* ---- fatal errors ---------------------------------------------------------- */

void
vl_host_fatal (char const * format, ...)
{
  char msg [1024] ;
  va_list args ;
  va_start(args, format) ;
  vsnprintf(msg, sizeof(msg), format, args) ;
  va_end(args) ;
  if (g_printf != quiet_printf) g_printf("%s", msg) ;
  fflush(stdout) ;
  fprintf(stderr, "%s", msg) ;
  fflush(stderr) ;
  abort() ;
}

/* ---- error state of the calling thread ------------------------------------ */

static __thread int g_last_error = VL_ERR_OK ;
static __thread char g_last_error_message [VL_ERR_MSG_LEN] ;

VL_EXPORT int
vl_set_last_error (int error, char const * errorMessage, ...)
{
  va_list args ;
  va_start(args, errorMessage) ;
  if (errorMessage) {
    vsnprintf(g_last_error_message, sizeof(g_last_error_message), errorMessage, args) ;
  } else {
    g_last_error_message[0] = 0 ;
  }
  va_end(args) ;
  g_last_error = error ;
  return error ;
}

VL_EXPORT int vl_get_last_error (void) { return g_last_error ; }
VL_EXPORT char const * vl_get_last_error_message (void) { return g_last_error_message ; }

/* ---- timers --------------------------------------------------------------- */

static __thread double g_tic = 0 ;

static double
now_seconds (void)
{
  struct timespec ts ;
  clock_gettime(CLOCK_MONOTONIC, &ts) ;
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec ;
}

VL_EXPORT void vl_tic (void) { g_tic = now_seconds() ; }
VL_EXPORT double vl_toc (void) { return now_seconds() - g_tic ; }

VL_EXPORT double
vl_get_cpu_time (void)
{
  return (double)clock() / (double)CLOCKS_PER_SEC ;
}
