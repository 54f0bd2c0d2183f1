// row_order.cpp -- a locality order of the rows of a sparse graph for the row kernels.
//
// The batched kernels gather K systems' values of every neighbour (K*B values per node, e.g.
// 384 bytes for 32 right-hand sides of a 3-component field in FP32).  In a lexicographic 3D
// numbering the rows that run at the same time on one XCD reach two planes up and down, a
// window larger than that XCD's L2, so x is fetched from memory several times per sweep.
// Processing the rows tile by tile -- compact regions of about `tile` rows grown by
// breadth-first search over the graph -- keeps the window small.  The order only changes
// which rows are evaluated together, never the arithmetic of a row, and needs no geometry, so
// it applies to any numbering (the reference's hierarchical refinement numbering included).

#include <algorithm>
#include <cstdint>
#include <vector>

typedef std::uint32_t u32;
typedef std::uint64_t u64;

extern "C" {

// order: a permutation of 0..rows-1; tiles are grown from the lowest unvisited row, rows of a
// tile are listed in ascending order, and tiles follow one another in seed order
int feat3_setup_tile_order(u64 rows, const u32* row_ptr, const u32* col_ind, u64 tile, u32* order)
{
  if(!order || (rows && !row_ptr)) return 1;
  if(tile <= 1)
  {
    for(u64 i = 0; i < rows; ++i) order[i] = u32(i);
    return 0;
  }
  for(u64 i = 0; i < rows; ++i)
    for(u32 e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
      if(u64(col_ind[e]) >= rows) return 1;
  std::vector<char> seen(rows, 0);
  std::vector<u32> queue, members;
  queue.reserve(tile * 4);
  u64 out = 0, seed = 0;
  while(out < rows)
  {
    while(seen[seed]) ++seed;
    queue.clear();
    members.clear();
    queue.push_back(u32(seed));
    seen[seed] = 1;
    std::size_t head = 0;
    while(head < queue.size() && members.size() < tile)
    {
      const u32 v = queue[head++];
      members.push_back(v);
      for(u32 e = row_ptr[v]; e < row_ptr[v + 1]; ++e)
      {
        const u32 w = col_ind[e];
        if(!seen[w]) { seen[w] = 1; queue.push_back(w); }
      }
    }
    // nodes queued but not taken return to the pool (and may lie below the next seed)
    for(std::size_t i = head; i < queue.size(); ++i)
    {
      seen[queue[i]] = 0;
      if(queue[i] < seed) seed = queue[i];
    }
    std::sort(members.begin(), members.end());
    for(u32 v : members) order[out++] = v;
  }
  return 0;
}

} // extern "C"
