"""Serving loop: requests (mesh id + camera + viewport) answered by batched
ray-casts over device-resident trees cached by mesh id (no reference
counterpart)."""
