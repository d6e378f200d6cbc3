// Base of the semantic landmarks (sloam/include/objects/semanticObject.h:6-18).
#pragma once

#include <cstddef>

#include "definitions.h"

template <typename T>
class SemanticObject {
 public:
  virtual ~SemanticObject() = default;
  virtual Scalar distance(const T& model) const = 0;
  virtual Scalar distance(const PointT& point) const = 0;
  virtual void project(const SE3& tf) = 0;
  T getModel() const { return model; }
  VectorType getFeatures() const { return features; }
  size_t id = 0;
  bool isValid = false;
  VectorType features;
  T model;
};
