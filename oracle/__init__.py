"""CPU oracle for pylbm's fused stream-collide step -- TEST INFRASTRUCTURE ONLY.

``lbm_ref`` is the float64 NumPy restatement of the reference algorithm (pinned against the
reference's golden vectors in tests/golden/); ``lbm_ref_c`` is a per-node C restatement of
the same step for larger checks; ``torch_ref`` is a float64 torch restatement for autograd
gradient checks.  The product package ``pylbm_amd`` never imports anything from here.
"""
