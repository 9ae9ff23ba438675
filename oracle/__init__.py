"""The oracle: a CPU restatement of the reference's read path, used only by tests, the smoke
check and the CPU-baseline leg of the benchmark, as the checker. Never part of the product path."""
