"""Algorithms on the GPU engine (reference ``fqe.algorithm``)."""
