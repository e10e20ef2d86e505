// Host-side steps of the NSG-style graph construction shared by the CPU
// builder (ss_build_graph) and the one whose prune runs on the GPU
// (ss_build_graph_gpu).  Not installed; plain C++ types only.
#pragma once

#include <cstddef>
#include <cstdint>

namespace ss {

// Distance with the search's arithmetic (eight partial sums over 16-float
// blocks with a fused multiply-add, pairwise final sum).
float host_distance(int metric, const float *a, const float *b, uint32_t dim);

// Connectivity pass on a graph held as rows of (id, distance) sorted by
// (distance, id): ids/dists [num][R], deg [num].  A node u not reachable
// from ep is attached to a target: its nearest (by (distance, id))
// reachable node with a free slot among its knn[u] and extra[u] lists, else
// among the nodes a best-first search for u from ep over the graph reaches
// (NSG's connectivity step), else among every reachable node.  The edge is
// inserted at its sorted place and the nodes u reaches become reachable.
// The unreachable nodes are taken in id order in batches: the targets of a
// batch are found in parallel on the graph as the batch starts, then
// applied in id order; a node reached meanwhile is skipped, and one whose
// target has filled up meanwhile is tried again, before the rest, in the
// next batch.  The result does not depend on the thread count.  Returns
// SS_OK, or SS_EINVAL when a node has nowhere to go; *fixed counts the
// attached nodes.
int connect_graph(const float *base, uint32_t num, uint32_t dim, int metric, const uint32_t *knn,
                  uint32_t knn_k, const uint32_t *extra, uint32_t extra_k, uint32_t R, uint32_t ep,
                  uint32_t *ids, float *dists, uint32_t *deg, uint32_t *fixed);

// The reverse-edge step of one node u on the host: the union of u's pruned
// list (ids/dists, n entries) and the nodes whose pruned lists hold u
// (rev, m ids), sorted by (distance, id); when it holds more than R
// entries it is pruned again by the occlusion rule with factor alpha.
// Writes at most R entries to out_ids/out_dists and their count to *out_deg.
void merge_reverse_host(const float *base, uint32_t dim, int metric, float alpha, uint32_t u, const uint32_t *ids,
                        const float *dists, size_t n, const uint32_t *rev, size_t m, uint32_t R, uint32_t *out_ids,
                        float *out_dists, uint32_t *out_deg);

}  // namespace ss
