// pagerankTypes.hxx -- PagerankOptions / PagerankResult for programs that do not include the reference's
// inc/pagerank.hxx. Same fields, defaults and constructors as pagerank.hxx:24-113, so code written against the
// reference's types compiles against these; include exactly one of the two definitions.
#pragma once
#include <utility>
#include <vector>

template <class V>
struct PagerankOptions {
  int repeat;
  V tolerance;
  V frontierTolerance;
  V pruneTolerance;
  V damping;
  int maxIterations;

  PagerankOptions(int repeat = 1, V tolerance = 1e-10, V frontierTolerance = 1e-6, V pruneTolerance = 1e-6,
                  V damping = 0.85, int maxIterations = 500)
      : repeat(repeat), tolerance(tolerance), frontierTolerance(frontierTolerance), pruneTolerance(pruneTolerance),
        damping(damping), maxIterations(maxIterations) {}
};

template <class V>
struct PagerankResult {
  std::vector<V> ranks;
  int iterations;
  float time;
  float initializationTime;
  float markingTime;
  float computationTime;

  PagerankResult() : ranks(), iterations(0), time(0), initializationTime(0), markingTime(0), computationTime(0) {}
  PagerankResult(std::vector<V>&& ranks, int iterations = 0, float time = 0, float initializationTime = 0,
                 float markingTime = 0, float computationTime = 0)
      : ranks(std::move(ranks)), iterations(iterations), time(time), initializationTime(initializationTime),
        markingTime(markingTime), computationTime(computationTime) {}
  PagerankResult(std::vector<V>& ranks, int iterations = 0, float time = 0, float initializationTime = 0,
                 float markingTime = 0, float computationTime = 0)
      : ranks(std::move(ranks)), iterations(iterations), time(time), initializationTime(initializationTime),
        markingTime(markingTime), computationTime(computationTime) {}
};
