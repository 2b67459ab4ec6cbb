/* kdtree.c -- randomised kd-forest (best-bin-first search), host side.
 *
 * Follows the algorithm of vl/kdtree.c (VLFeat 0.9.21):
 *  - build (kdtree.c:165-305, 530-580): each node with more than one point
 *    estimates the variance of every dimension (all points when the node holds
 *    at most 1024, else 1024 random samples drawn with replacement from the
 *    first 1024 entries), keeps the min(numTrees, 5) most varying dimensions
 *    in a binary min-heap, draws one generator word to pick a heap slot, sorts
 *    the node's points along that dimension with the C library qsort and
 *    splits at the median;
 *  - bounds (kdtree.c:497-521): every node records the interval of its split
 *    dimension inherited from its ancestors;
 *  - query (kdtree.c:585-847): branch and bound over a priority queue of
 *    partially explored subtrees of all trees, each data point compared at
 *    most once per query, stopping after maxNumComparisons comparisons (0 =
 *    exact search) or when no queued subtree can hold a closer point.
 * The three binary heaps (split dimensions, search states, neighbours) use the
 * sift rules of vl/heap-def.h so ties resolve in the same order.
 */
#include "vl/kdtree.h"

#include <assert.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  vl_uindex parent ;
  vl_index lowerChild ;
  vl_index upperChild ;
  unsigned int splitDimension ;
  double splitThreshold ;
  double lowerBound ;
  double upperBound ;
} KdNode ;

typedef struct {
  vl_index index ;
  double value ;
} KdEntry ;

typedef struct {
  KdNode * nodes ;
  vl_size numUsedNodes ;
  vl_size numAllocatedNodes ;
  KdEntry * dataIndex ;
  unsigned int depth ;
} KdTree ;

typedef struct {
  unsigned int dimension ;
  double mean ;
  double variance ;
} KdSplit ;

typedef struct {
  KdTree * tree ;
  vl_uindex nodeIndex ;
  double distanceLowerBound ;
} KdState ;

struct _VlKDForest {
  vl_type dataType ;
  vl_size dimension ;
  vl_size numData ;
  void const * data ;
  vl_size numTrees ;
  KdTree ** trees ;
  VlKDTreeThresholdingMethod thresholdingMethod ;
  KdSplit splitHeap [VL_KDTREE_SPLIT_HEAP_SIZE] ;
  vl_size splitHeapSize ;
  vl_size splitHeapNum ;
  vl_size maxNumNodes ;
  vl_size maxNumComparisons ;
  VlRand * rand ;
  VlFloatVectorComparisonFunction distF ;
  VlDoubleVectorComparisonFunction distD ;
  VlKDForestSearcher * searchers ;
} ;

struct _VlKDForestSearcher {
  VlKDForest * forest ;
  KdState * heap ;
  vl_size heapNum ;
  vl_uindex * idBook ;
  vl_uindex searchId ;
  vl_size numComparisons ;
  VlKDForestSearcher * next ;
} ;

/* ---- binary heaps (root = smallest under the given order) -------------- */

#define DEFINE_HEAP(PFX, T, CMP)                                               \
  static void PFX##_swap (T * a, vl_uindex i, vl_uindex j) { T t = a[i] ; a[i] = a[j] ; a[j] = t ; } \
  static void PFX##_to_root (T * a, vl_uindex i) {                             \
    while (i > 0) {                                                            \
      vl_uindex p = (i - 1) / 2 ;                                              \
      if (!(CMP(a, i, p) < 0)) return ;                                        \
      PFX##_swap(a, i, p) ; i = p ;                                            \
    }                                                                          \
  }                                                                            \
  static void PFX##_to_leaves (T * a, vl_size n, vl_uindex i) {                \
    for (;;) {                                                                 \
      vl_uindex l = 2 * i + 1, r = 2 * i + 2 ;                                 \
      if (l >= n) return ;                                                     \
      if (r >= n) { if (CMP(a, i, l) > 0) PFX##_swap(a, i, l) ; return ; }     \
      if (CMP(a, l, r) < 0) {                                                  \
        if (!(CMP(a, i, l) > 0)) return ;                                      \
        PFX##_swap(a, i, l) ; i = l ;                                          \
      } else {                                                                 \
        if (!(CMP(a, i, r) > 0)) return ;                                      \
        PFX##_swap(a, i, r) ; i = r ;                                          \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  static void PFX##_push (T * a, vl_size * n) { PFX##_to_root(a, *n) ; *n += 1 ; } \
  static vl_uindex PFX##_pop (T * a, vl_size * n) {                            \
    *n -= 1 ; PFX##_swap(a, 0, *n) ; if (*n > 1) PFX##_to_leaves(a, *n, 0) ; return *n ; \
  }                                                                            \
  static void PFX##_update (T * a, vl_size n, vl_uindex i) { PFX##_to_leaves(a, n, i) ; PFX##_to_root(a, i) ; }

#define CMP_SPLIT(a, x, y) ((a)[x].variance - (a)[y].variance)
#define CMP_STATE(a, x, y) ((a)[x].distanceLowerBound - (a)[y].distanceLowerBound)
#define CMP_NEIGH(a, x, y) ((a)[y].distance - (a)[x].distance)

DEFINE_HEAP(split_heap, KdSplit, CMP_SPLIT)
DEFINE_HEAP(state_heap, KdState, CMP_STATE)
DEFINE_HEAP(neigh_heap, VlKDForestNeighbor, CMP_NEIGH)

/* ---- construction -------------------------------------------------------- */

static int
entry_compare (void const * a, void const * b)
{
  double delta = ((KdEntry const*)a)->value - ((KdEntry const*)b)->value ;
  if (delta < 0) return -1 ;
  if (delta > 0) return +1 ;
  return 0 ;
}

static vl_uindex
node_new (KdTree * tree, vl_uindex parent)
{
  vl_uindex i = tree->numUsedNodes ++ ;
  KdNode * n ;
  assert(tree->numUsedNodes <= tree->numAllocatedNodes) ;
  n = tree->nodes + i ;
  n->parent = parent ;
  n->lowerChild = 0 ;
  n->upperChild = 0 ;
  n->splitDimension = 0 ;
  n->splitThreshold = 0 ;
  return i ;
}

static double
datum (VlKDForest const * f, vl_index i, vl_uindex d)
{
  if (f->dataType == VL_TYPE_FLOAT) return ((float const*)f->data)[i * f->dimension + d] ;
  return ((double const*)f->data)[i * f->dimension + d] ;
}

static void
build_node (VlKDForest * f, KdTree * tree, vl_uindex nodeIndex,
            vl_uindex begin, vl_uindex end, unsigned int depth)
{
  KdNode * node = tree->nodes + nodeIndex ;
  KdSplit * split ;
  vl_uindex d, i, splitIndex ;

  if (end - begin <= 1) {
    if (tree->depth < depth) tree->depth = depth ;
    node->lowerChild = - (vl_index)begin - 1 ;
    node->upperChild = - (vl_index)end - 1 ;
    return ;
  }

  f->splitHeapNum = 0 ;
  for (d = 0 ; d < f->dimension ; ++d) {
    double mean = 0, second = 0, variance ;
    vl_size numSamples = VL_KDTREE_VARIANCE_EST_NUM_SAMPLES ;
    vl_bool all = VL_FALSE ;
    if (end - begin <= VL_KDTREE_VARIANCE_EST_NUM_SAMPLES) {
      all = VL_TRUE ;
      numSamples = end - begin ;
    }
    for (i = 0 ; i < numSamples ; ++i) {
      vl_uint32 s = all ? (vl_uint32)i : (vl_rand_uint32(f->rand) % VL_KDTREE_VARIANCE_EST_NUM_SAMPLES) ;
      double x = datum(f, tree->dataIndex[begin + s].index, d) ;
      mean += x ;
      second += x * x ;
    }
    mean /= numSamples ;
    second /= numSamples ;
    variance = second - mean * mean ;
    if (variance <= 0) continue ;
    if (f->splitHeapNum < f->splitHeapSize) {
      KdSplit * s = f->splitHeap + f->splitHeapNum ;
      s->dimension = (unsigned int)d ;
      s->mean = mean ;
      s->variance = variance ;
      split_heap_push(f->splitHeap, &f->splitHeapNum) ;
    } else if (f->splitHeap[0].variance < variance) {
      f->splitHeap[0].dimension = (unsigned int)d ;
      f->splitHeap[0].mean = mean ;
      f->splitHeap[0].variance = variance ;
      split_heap_update(f->splitHeap, f->splitHeapNum, 0) ;
    }
  }

  if (f->splitHeapNum == 0) {
    node->lowerChild = - (vl_index)begin - 1 ;
    node->upperChild = - (vl_index)end - 1 ;
    return ;
  }

  split = f->splitHeap + (vl_rand_uint32(f->rand) % VL_MIN(f->splitHeapSize, f->splitHeapNum)) ;
  node->splitDimension = split->dimension ;
  for (i = begin ; i < end ; ++i) {
    tree->dataIndex[i].value = datum(f, tree->dataIndex[i].index, split->dimension) ;
  }
  qsort(tree->dataIndex + begin, end - begin, sizeof(KdEntry), entry_compare) ;

  splitIndex = (begin + end - 1) / 2 ;
  node->splitThreshold = tree->dataIndex[splitIndex].value ;
  if (f->thresholdingMethod == VL_KDTREE_MEAN) {
    vl_uindex s ;
    double t = split->mean ;
    for (s = begin ; s < end && tree->dataIndex[s].value <= t ; ++s) ;
    if (s >= begin + 1 && s < end) {          /* proper partition by the mean */
      splitIndex = s - 1 ;
      node->splitThreshold = t ;
    }
  }

  {
    vl_uindex lower = node_new(tree, nodeIndex) ;
    tree->nodes[nodeIndex].lowerChild = (vl_index)lower ;
    build_node(f, tree, lower, begin, splitIndex + 1, depth + 1) ;
  }
  {
    vl_uindex upper = node_new(tree, nodeIndex) ;
    tree->nodes[nodeIndex].upperChild = (vl_index)upper ;
    build_node(f, tree, upper, splitIndex + 1, end, depth + 1) ;
  }
}

static void
calc_bounds (KdTree * tree, vl_uindex nodeIndex, double * bounds)
{
  KdNode * node = tree->nodes + nodeIndex ;
  vl_uindex i = node->splitDimension ;
  double t = node->splitThreshold ;
  node->lowerBound = bounds[2 * i + 0] ;
  node->upperBound = bounds[2 * i + 1] ;
  if (node->lowerChild > 0) {
    bounds[2 * i + 1] = t ;
    calc_bounds(tree, (vl_uindex)node->lowerChild, bounds) ;
    bounds[2 * i + 1] = node->upperBound ;
  }
  if (node->upperChild > 0) {
    bounds[2 * i + 0] = t ;
    calc_bounds(tree, (vl_uindex)node->upperChild, bounds) ;
    bounds[2 * i + 0] = node->lowerBound ;
  }
}

VL_EXPORT VlKDForest *
vl_kdforest_new (vl_type dataType, vl_size dimension, vl_size numTrees, VlVectorComparisonType normType)
{
  VlKDForest * self = (VlKDForest*) vl_calloc(1, sizeof(VlKDForest)) ;
  assert(dataType == VL_TYPE_FLOAT || dataType == VL_TYPE_DOUBLE) ;
  assert(dimension >= 1) ;
  assert(numTrees >= 1) ;
  self->rand = vl_get_rand() ;
  self->dataType = dataType ;
  self->dimension = dimension ;
  self->numTrees = numTrees ;
  self->thresholdingMethod = VL_KDTREE_MEDIAN ;
  self->splitHeapSize = VL_MIN(numTrees, VL_KDTREE_SPLIT_HEAP_SIZE) ;
  if (dataType == VL_TYPE_FLOAT) self->distF = vl_get_vector_comparison_function_f(normType) ;
  else self->distD = vl_get_vector_comparison_function_d(normType) ;
  return self ;
}

VL_EXPORT void
vl_kdforest_build (VlKDForest * self, vl_size numData, void const * data)
{
  vl_uindex ti, di ;
  double * bounds ;
  assert(data) ;
  assert(numData >= 1) ;
  self->data = data ;
  self->numData = numData ;
  self->trees = (KdTree**) vl_malloc(sizeof(KdTree*) * self->numTrees) ;
  self->maxNumNodes = 0 ;
  for (ti = 0 ; ti < self->numTrees ; ++ti) {
    KdTree * t = (KdTree*) vl_malloc(sizeof(KdTree)) ;
    t->dataIndex = (KdEntry*) vl_malloc(sizeof(KdEntry) * numData) ;
    for (di = 0 ; di < numData ; ++di) { t->dataIndex[di].index = (vl_index)di ; t->dataIndex[di].value = 0 ; }
    t->numUsedNodes = 0 ;
    t->numAllocatedNodes = 2 * numData - 1 ;
    t->nodes = (KdNode*) vl_malloc(sizeof(KdNode) * t->numAllocatedNodes) ;
    t->depth = 0 ;
    self->trees[ti] = t ;
    build_node(self, t, node_new(t, 0), 0, numData, 0) ;
    self->maxNumNodes += t->numUsedNodes ;
  }
  bounds = (double*) vl_malloc(sizeof(double) * 2 * self->dimension) ;
  for (ti = 0 ; ti < self->numTrees ; ++ti) {
    vl_uindex d ;
    for (d = 0 ; d < self->dimension ; ++d) {
      bounds[2 * d + 0] = -VL_INFINITY_F ;
      bounds[2 * d + 1] = +VL_INFINITY_F ;
    }
    calc_bounds(self->trees[ti], 0, bounds) ;
  }
  vl_free(bounds) ;
}

VL_EXPORT void vl_kdforest_set_max_num_comparisons (VlKDForest * self, vl_size n) { self->maxNumComparisons = n ; }
VL_EXPORT vl_size vl_kdforest_get_max_num_comparisons (VlKDForest * self) { return self->maxNumComparisons ; }
VL_EXPORT void vl_kdforest_set_thresholding_method (VlKDForest * self, VlKDTreeThresholdingMethod m) { self->thresholdingMethod = m ; }
VL_EXPORT vl_size vl_kdforest_get_num_nodes_of_tree (VlKDForest const * self, vl_uindex t) { return self->trees[t]->numUsedNodes ; }
VL_EXPORT vl_size vl_kdforest_get_depth_of_tree (VlKDForest const * self, vl_uindex t) { return self->trees[t]->depth ; }

VL_EXPORT VlKDForestSearcher *
vl_kdforest_new_searcher (VlKDForest * f)
{
  VlKDForestSearcher * s = (VlKDForestSearcher*) vl_calloc(1, sizeof(VlKDForestSearcher)) ;
  s->forest = f ;
  s->heap = (KdState*) vl_malloc(sizeof(KdState) * (f->maxNumNodes ? f->maxNumNodes : 1)) ;
  s->idBook = (vl_uindex*) vl_calloc(f->numData ? f->numData : 1, sizeof(vl_uindex)) ;
  /* appended: searchers are numbered in creation order (kdtree.c:387-407) */
  {
    VlKDForestSearcher ** p = &f->searchers ;
    while (*p) p = &(*p)->next ;
    s->next = NULL ;
    *p = s ;
  }
  return s ;
}

VL_EXPORT void
vl_kdforestsearcher_delete (VlKDForestSearcher * s)
{
  VlKDForestSearcher ** p = &s->forest->searchers ;
  while (*p && *p != s) p = &(*p)->next ;
  if (*p) *p = s->next ;
  vl_free(s->heap) ;
  vl_free(s->idBook) ;
  vl_free(s) ;
}

VL_EXPORT void
vl_kdforest_delete (VlKDForest * self)
{
  vl_uindex ti ;
  while (self->searchers) vl_kdforestsearcher_delete(self->searchers) ;
  if (self->trees) {
    for (ti = 0 ; ti < self->numTrees ; ++ti) {
      if (self->trees[ti]) {
        vl_free(self->trees[ti]->nodes) ;
        vl_free(self->trees[ti]->dataIndex) ;
        vl_free(self->trees[ti]) ;
      }
    }
    vl_free(self->trees) ;
  }
  vl_free(self) ;
}

/* ---- query --------------------------------------------------------------- */

static void
query_node (VlKDForestSearcher * s, KdTree * tree, vl_uindex nodeIndex,
            VlKDForestNeighbor * nb, vl_size numNb, vl_size * numAdded,
            double dist, void const * query)
{
  VlKDForest * f = s->forest ;
  for (;;) {
    KdNode const * node = tree->nodes + nodeIndex ;
    vl_uindex i = node->splitDimension ;
    double x = (f->dataType == VL_TYPE_FLOAT) ? ((float const*)query)[i] : ((double const*)query)[i] ;
    double x1 = node->lowerBound, x2 = node->splitThreshold, x3 = node->upperBound ;
    double delta, saveDist ;
    vl_index nextChild, saveChild ;

    if (node->lowerChild < 0) {
      vl_index it, b = - node->lowerChild - 1, e = - node->upperChild - 1 ;
      for (it = b ; it < e && (f->maxNumComparisons == 0 || s->numComparisons < f->maxNumComparisons) ; ++it) {
        vl_index di = tree->dataIndex[it].index ;
        double dd ;
        if (s->idBook[di] == s->searchId) continue ;
        s->idBook[di] = s->searchId ;
        if (f->dataType == VL_TYPE_FLOAT) {
          dd = f->distF(f->dimension, (float const*)query, (float const*)f->data + di * f->dimension) ;
        } else {
          dd = f->distD(f->dimension, (double const*)query, (double const*)f->data + di * f->dimension) ;
        }
        s->numComparisons += 1 ;
        if (*numAdded < numNb) {
          nb[*numAdded].index = (vl_uindex)di ;
          nb[*numAdded].distance = dd ;
          neigh_heap_push(nb, numAdded) ;
        } else if (nb[0].distance > dd) {
          nb[0].index = (vl_uindex)di ;
          nb[0].distance = dd ;
          neigh_heap_update(nb, *numAdded, 0) ;
        }
      }
      return ;
    }

    delta = x - x2 ;
    saveDist = dist + delta * delta ;
    if (x <= x2) {
      nextChild = node->lowerChild ;
      saveChild = node->upperChild ;
      if (x <= x1) { delta = x - x1 ; saveDist -= delta * delta ; }
    } else {
      nextChild = node->upperChild ;
      saveChild = node->lowerChild ;
      if (x > x3) { delta = x - x3 ; saveDist -= delta * delta ; }
    }
    if (*numAdded < numNb || nb[0].distance > saveDist) {
      KdState * st = s->heap + s->heapNum ;
      st->tree = tree ;
      st->nodeIndex = (vl_uindex)saveChild ;
      st->distanceLowerBound = saveDist ;
      state_heap_push(s->heap, &s->heapNum) ;
    }
    nodeIndex = (vl_uindex)nextChild ;
  }
}

VL_EXPORT vl_size
vl_kdforestsearcher_query (VlKDForestSearcher * s, VlKDForestNeighbor * nb,
                           vl_size numNb, void const * query)
{
  VlKDForest * f = s->forest ;
  vl_bool exact = (f->maxNumComparisons == 0) ;
  vl_size numAdded = 0 ;
  vl_uindex i, ti ;
  assert(nb && numNb > 0 && query) ;
  s->searchId += 1 ;
  s->numComparisons = 0 ;
  s->heapNum = 0 ;
  for (ti = 0 ; ti < f->numTrees ; ++ti) {
    KdState * st = s->heap + s->heapNum ;
    st->tree = f->trees[ti] ;
    st->nodeIndex = 0 ;
    st->distanceLowerBound = 0 ;
    state_heap_push(s->heap, &s->heapNum) ;
  }
  while (exact || s->numComparisons < f->maxNumComparisons) {
    KdState * st ;
    if (s->heapNum == 0) break ;
    st = s->heap + state_heap_pop(s->heap, &s->heapNum) ;
    if (numAdded == numNb && nb[0].distance < st->distanceLowerBound) break ;
    query_node(s, st->tree, st->nodeIndex, nb, numNb, &numAdded, st->distanceLowerBound, query) ;
  }
  for (i = numAdded ; i < numNb ; ++i) {
    nb[i].index = (vl_uindex)-1 ;
    nb[i].distance = VL_NAN_F ;
  }
  while (numAdded) neigh_heap_pop(nb, &numAdded) ;
  return s->numComparisons ;
}

VL_EXPORT vl_size
vl_kdforest_query (VlKDForest * self, VlKDForestNeighbor * nb, vl_size numNb, void const * query)
{
  VlKDForestSearcher * s = self->searchers ? self->searchers : vl_kdforest_new_searcher(self) ;
  return vl_kdforestsearcher_query(s, nb, numNb, query) ;
}

/* kdtree.c:870-946: numQueries queries at once, the queries split over
 * vl_get_max_threads() host threads, each with its own searcher; indexes and
 * distances are numNeighbors x numQueries (distances in the data type, may be
 * NULL).  Returns the total number of comparisons. */
VL_EXPORT vl_size
vl_kdforest_query_with_array (VlKDForest * self, vl_uint32 * indexes, vl_size numNeighbors,
                              vl_size numQueries, void * distances, void const * queries)
{
  vl_size total = 0 ;
  vl_size const dim = self->dimension ;
  vl_type const type = self->dataType ;
  if (type != VL_TYPE_FLOAT && type != VL_TYPE_DOUBLE) abort() ;
#pragma omp parallel num_threads((int)vl_get_max_threads()) reduction(+:total)
  {
    VlKDForestSearcher * s ;
    VlKDForestNeighbor * nb ;
    vl_index q ;
#pragma omp critical (vl_kdforest_searchers)
    {
      s = vl_kdforest_new_searcher(self) ;
      nb = (VlKDForestNeighbor*) vl_calloc(numNeighbors ? numNeighbors : 1, sizeof(VlKDForestNeighbor)) ;
    }
#pragma omp for
    for (q = 0 ; q < (vl_index)numQueries ; ++q) {
      vl_uindex n ;
      void const * query = (type == VL_TYPE_FLOAT) ? (void const*)((float const*)queries + q * dim)
                                                    : (void const*)((double const*)queries + q * dim) ;
      total += vl_kdforestsearcher_query(s, nb, numNeighbors, query) ;
      for (n = 0 ; n < numNeighbors ; ++n) {
        indexes[q * numNeighbors + n] = (vl_uint32) nb[n].index ;
        if (distances) {
          if (type == VL_TYPE_FLOAT) ((float*)distances)[q * numNeighbors + n] = (float) nb[n].distance ;
          else ((double*)distances)[q * numNeighbors + n] = nb[n].distance ;
        }
      }
    }
#pragma omp critical (vl_kdforest_searchers)
    {
      vl_kdforestsearcher_delete(s) ;
      vl_free(nb) ;
    }
  }
  return total ;
}

VL_EXPORT vl_size vl_kdforest_get_num_trees (VlKDForest const * self) { return self->numTrees ; }
VL_EXPORT vl_size vl_kdforest_get_data_dimension (VlKDForest const * self) { return self->dimension ; }
VL_EXPORT vl_type vl_kdforest_get_data_type (VlKDForest const * self) { return self->dataType ; }
VL_EXPORT VlKDTreeThresholdingMethod
vl_kdforest_get_thresholding_method (VlKDForest const * self) { return self->thresholdingMethod ; }
VL_EXPORT VlKDForest * vl_kdforest_searcher_get_forest (VlKDForestSearcher const * self) { return self->forest ; }

/* the pos-th searcher in creation order, NULL past the last (kdtree.c:443-452) */
VL_EXPORT VlKDForestSearcher *
vl_kdforest_get_searcher (VlKDForest const * self, vl_uindex pos)
{
  VlKDForestSearcher * s = self->searchers ;
  while (s && pos) { s = s->next ; --pos ; }
  return s ;
}
