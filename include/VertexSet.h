// VertexSet — sorted vertex-id set with the CPU set operations of the
// reference's VertexSet (include/VertexSet.h:21-268, common/VertexSet.cc:21-89):
// merge intersection (count / materialise, optional strict upper bound and
// excluded ancestors), difference that drops the owner vertex of the
// subtrahend (`other.vid`), and bounded prefixes.  Host-only; the device
// kernels of libjgpm do the same operations on the GPU.
#pragma once
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

typedef int32_t vidType;
typedef int64_t eidType;
typedef unsigned long long AccType;

class VertexSet {
 public:
  static vidType MAX_DEGREE;  // capacity of owning sets
  static void release_buffers() {}

  VertexSet() : ptr_(nullptr), size_(0), vid_(-1) {
    own_.resize((size_t)MAX_DEGREE);
    ptr_ = own_.data();
  }
  // non-owning view of a neighbour list (Graph::N)
  VertexSet(vidType* p, vidType s, vidType id) : ptr_(p), size_(s), vid_(id) {}
  VertexSet(const VertexSet&) = delete;
  VertexSet& operator=(const VertexSet&) = delete;
  VertexSet(VertexSet&& o) noexcept { *this = std::move(o); }
  VertexSet& operator=(VertexSet&& o) noexcept {
    own_ = std::move(o.own_);
    ptr_ = own_.empty() ? o.ptr_ : own_.data();
    size_ = o.size_;
    vid_ = o.vid_;
    return *this;
  }

  vidType size() const { return size_; }
  vidType get_vid() const { return vid_; }
  vidType* data() const { return ptr_; }
  const vidType* begin() const { return ptr_; }
  const vidType* end() const { return ptr_ + size_; }
  void add(vidType v) { ptr_[size_++] = v; }
  void clear() { size_ = 0; }
  vidType& operator[](size_t i) { return ptr_[i]; }
  const vidType& operator[](size_t i) const { return ptr_[i]; }

  VertexSet operator&(const VertexSet& other) const;  // intersection
  uint32_t get_intersect_num(const VertexSet& other) const;
  VertexSet intersect(const VertexSet& other, vidType upper) const;      // elements < upper
  vidType intersect_ns(const VertexSet& other, vidType upper) const;    // count, elements < upper
  vidType intersect_ns_except(const VertexSet& other, vidType upper, vidType ancestor) const;
  vidType intersect_except(const VertexSet& other, vidType ancestor) const;
  vidType intersect_except(const VertexSet& other, vidType ancestorA, vidType ancestorB) const;

  // this - other, also dropping other's owner vertex; outBuf may alias ptr
  vidType difference_buf(vidType* outBuf, const VertexSet& other) const;
  vidType difference_buf(vidType* outBuf, const VertexSet& other, vidType upper) const;
  VertexSet operator-(const VertexSet& other) const;
  VertexSet& difference(VertexSet& dst, const VertexSet& other) const;
  VertexSet difference(const VertexSet& other, vidType upper) const;
  VertexSet& difference(VertexSet& dst, const VertexSet& other, vidType upper) const;
  vidType difference_ns(const VertexSet& other, vidType upper) const;

  VertexSet bounded(vidType up) const;  // prefix of elements < up (a view)

 private:
  std::vector<vidType> own_;
  vidType* ptr_;
  vidType size_;
  vidType vid_;
};

inline VertexSet difference_set(const VertexSet& a, const VertexSet& b) { return a - b; }
inline VertexSet& difference_set(VertexSet& dst, const VertexSet& a, const VertexSet& b) { return a.difference(dst, b); }
inline VertexSet difference_set(const VertexSet& a, const VertexSet& b, vidType up) { return a.difference(b, up); }
inline VertexSet& difference_set(VertexSet& dst, const VertexSet& a, const VertexSet& b, vidType up) {
  return a.difference(dst, b, up);
}
inline uint64_t difference_num(const VertexSet& a, const VertexSet& b) { return (uint64_t)(a - b).size(); }
inline uint64_t difference_num(const VertexSet& a, const VertexSet& b, vidType up) { return a.difference_ns(b, up); }
inline VertexSet intersection_set(const VertexSet& a, const VertexSet& b) { return a & b; }
inline VertexSet intersection_set(const VertexSet& a, const VertexSet& b, vidType up) { return a.intersect(b, up); }
inline uint64_t intersection_num(const VertexSet& a, const VertexSet& b) { return a.get_intersect_num(b); }
inline uint64_t intersection_num(const VertexSet& a, const VertexSet& b, vidType up) { return a.intersect_ns(b, up); }
inline uint64_t intersection_num_except(const VertexSet& a, const VertexSet& b, vidType anc) {
  return a.intersect_except(b, anc);
}
inline uint64_t intersection_num_except(const VertexSet& a, const VertexSet& b, vidType ancA, vidType ancB) {
  return a.intersect_except(b, ancA, ancB);
}
inline uint64_t intersection_num_bound_except(const VertexSet& a, const VertexSet& b, vidType up, vidType anc) {
  return a.intersect_ns_except(b, up, anc);
}
inline VertexSet bounded(const VertexSet& a, vidType up) { return a.bounded(up); }
void intersection(const VertexSet& a, const VertexSet& b, VertexSet& c);
