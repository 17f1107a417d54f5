# SANKHYA solution file
# generated-by sih26_amd libsk_gpu (MI355X PDHG engine)
# All numbers carry 17 significant digits and round-trip exactly.
model closed_sum
source 
sense minimize
status optimal
algorithm pdhg-cpu
objective 2.0000000000000218
dual_bound 1.9999999999999984
mip_relative_gap 0.0001
mip_absolute_gap 9.9999999999999995e-07
objective_offset 0
certificate none
rows 1
columns 2
iterations 73
nodes 0
solve_seconds 0
primal_infeasibility 0
dual_infeasibility 0
integrality_violation 0

# name value reduced_cost basis_status
begin columns 2
C0 1.0000000000000109 7.7715611723760958e-16 unknown
C1 1.0000000000000109 7.7715611723760958e-16 unknown
end columns

# name activity dual basis_status
begin rows 1
R0 2.0000000000000218 -0.99999999999999922 unknown
end rows
